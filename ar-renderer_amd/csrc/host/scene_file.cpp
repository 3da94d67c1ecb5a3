// Scene files for the GPU path tracer: the .prscene.json format of the reference loader
// (src/scene/scene_loader.cpp:48-1317), OBJ / PLY meshes (src/geometry/mesh_io.cpp) and Mitsuba .vol
// grids (src/media/medium.cpp:130-166), read straight into the flat SceneDescEx that the reference's
// GPU exporter (src/integrators/gpu_path.cpp:104-456) builds from the loaded scene.
//
// What the exporter approximates or drops is approximated or dropped here, with the same warning
// (checker textures keep both colours and the u scale; roughness textures become the default).
// Distant, spot, IES and sun lights, image environments, the Preetham sky of sun-sky lights, image
// textures as a material's reflectance, and normal and bump maps (images in PFM, PPM, PNG, HDR or
// EXR), which the v4 exporter averages or skips, are kept: the extension has them.
// Everything the renderer cannot express and the exporter has no fallback for (image textures in
// other slots, noise colour textures, orthographic and realistic cameras, glTF / pbrt imports, noise
// media) is a load error naming the item. Illuminant RGB spectra (sigmoid x D65) are tabulated every
// 5 nm, as the exporter tabulates every spectrum.
#include "image_codecs.h"
#include "prender_host.h"
#include "scene_build.h"

#include <algorithm>
#include <charconv>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <functional>
#include <map>
#include <memory>
#include <set>
#include <sstream>
#include <stdexcept>
#include <string>
#include <tuple>
#include <unordered_map>
#include <vector>

namespace fs = std::filesystem;

namespace phost {
namespace {

struct LoadError : std::runtime_error {
    using std::runtime_error::runtime_error;
};
[[noreturn]] void Fail(const std::string &m) { throw LoadError(m); }

// ---- JSON with comments ----------------------------------------------------------------------
// Objects keep their keys sorted (std::map), the iteration order of the reference's nlohmann::json:
// lights are created, and the first camera is chosen, in key order.
struct J {
    enum Kind { Null, Bool, Num, Str, Arr, Obj } k = Null;
    bool b = false;
    double n = 0;
    std::string s;
    std::vector<J> a;
    std::map<std::string, J> o;

    bool isNum() const { return k == Num; }
    bool isStr() const { return k == Str; }
    bool isArr() const { return k == Arr; }
    bool isObj() const { return k == Obj; }
    bool has(const std::string &key) const { return k == Obj && o.count(key) != 0; }
    const J &at(const std::string &key) const {
        static const J null;
        if (k != Obj) return null;
        auto it = o.find(key);
        return it == o.end() ? null : it->second;
    }
    J &set(const std::string &key) {
        if (k != Obj) {
            *this = J();
            k = Obj;
        }
        return o[key];
    }
    size_t size() const { return k == Arr ? a.size() : k == Obj ? o.size() : 0; }
    static J num(double v) {
        J j;
        j.k = Num;
        j.n = v;
        return j;
    }
    static J str(const std::string &v) {
        J j;
        j.k = Str;
        j.s = v;
        return j;
    }
    static J boolean(bool v) {
        J j;
        j.k = Bool;
        j.b = v;
        return j;
    }
    static J object() {
        J j;
        j.k = Obj;
        return j;
    }
    static J array() {
        J j;
        j.k = Arr;
        return j;
    }
};

class JsonReader {
  public:
    JsonReader(const std::string &text, std::string name) : c_(text.data()), e_(text.data() + text.size()), b_(c_), name_(std::move(name)) {}
    J Document() {
        J v = Value();
        Skip();
        if (c_ != e_) Error("unexpected content after the document");
        return v;
    }

  private:
    [[noreturn]] void Error(const std::string &m) const {
        int line = 1 + int(std::count(b_, c_, '\n'));
        Fail(name_ + ":" + std::to_string(line) + ": " + m);
    }
    void Skip() {  // whitespace, // line comments and /* block */ comments
        for (;;) {
            while (c_ < e_ && (*c_ == ' ' || *c_ == '\t' || *c_ == '\n' || *c_ == '\r')) ++c_;
            if (c_ + 1 < e_ && c_[0] == '/' && c_[1] == '/') {
                while (c_ < e_ && *c_ != '\n') ++c_;
            } else if (c_ + 1 < e_ && c_[0] == '/' && c_[1] == '*') {
                const char *end = std::strstr(c_ + 2, "*/");
                if (!end || end >= e_) Error("unterminated comment");
                c_ = end + 2;
            } else {
                return;
            }
        }
    }
    bool Lit(const char *w) {
        size_t n = std::strlen(w);
        if (size_t(e_ - c_) >= n && std::memcmp(c_, w, n) == 0) {
            c_ += n;
            return true;
        }
        return false;
    }
    static void PutUtf8(std::string &s, uint32_t cp) {
        if (cp < 0x80) s += char(cp);
        else if (cp < 0x800) s += char(0xc0 | (cp >> 6)), s += char(0x80 | (cp & 0x3f));
        else if (cp < 0x10000) s += char(0xe0 | (cp >> 12)), s += char(0x80 | ((cp >> 6) & 0x3f)), s += char(0x80 | (cp & 0x3f));
        else
            s += char(0xf0 | (cp >> 18)), s += char(0x80 | ((cp >> 12) & 0x3f)), s += char(0x80 | ((cp >> 6) & 0x3f)),
                s += char(0x80 | (cp & 0x3f));
    }
    uint32_t Hex4() {
        if (e_ - c_ < 4) Error("bad \\u escape");
        uint32_t v = 0;
        for (int i = 0; i < 4; ++i) {
            char h = *c_++;
            v <<= 4;
            if (h >= '0' && h <= '9') v |= uint32_t(h - '0');
            else if (h >= 'a' && h <= 'f') v |= uint32_t(h - 'a' + 10);
            else if (h >= 'A' && h <= 'F') v |= uint32_t(h - 'A' + 10);
            else Error("bad \\u escape");
        }
        return v;
    }
    std::string String() {
        ++c_;  // opening quote
        std::string s;
        while (c_ < e_ && *c_ != '"') {
            char ch = *c_++;
            if (ch != '\\') {
                if ((unsigned char)ch < 0x20) Error("control character in string");
                s += ch;
                continue;
            }
            if (c_ >= e_) break;
            char esc = *c_++;
            switch (esc) {
            case '"': s += '"'; break;
            case '\\': s += '\\'; break;
            case '/': s += '/'; break;
            case 'b': s += '\b'; break;
            case 'f': s += '\f'; break;
            case 'n': s += '\n'; break;
            case 'r': s += '\r'; break;
            case 't': s += '\t'; break;
            case 'u': {
                uint32_t cp = Hex4();
                if (cp >= 0xd800 && cp < 0xdc00 && c_ + 1 < e_ && c_[0] == '\\' && c_[1] == 'u') {
                    c_ += 2;
                    uint32_t lo = Hex4();
                    cp = 0x10000 + ((cp - 0xd800) << 10) + (lo - 0xdc00);
                }
                PutUtf8(s, cp);
                break;
            }
            default: Error("bad escape in string");
            }
        }
        if (c_ >= e_) Error("unterminated string");
        ++c_;
        return s;
    }
    J Value() {
        Skip();
        if (c_ >= e_) Error("unexpected end of input");
        J v;
        if (*c_ == '{') {
            ++c_;
            v.k = J::Obj;
            Skip();
            if (c_ < e_ && *c_ == '}') {
                ++c_;
                return v;
            }
            for (;;) {
                Skip();
                if (c_ >= e_ || *c_ != '"') Error("expected a key");
                std::string key = String();
                Skip();
                if (c_ >= e_ || *c_ != ':') Error("expected ':'");
                ++c_;
                v.o[key] = Value();
                Skip();
                if (c_ < e_ && *c_ == ',') {
                    ++c_;
                    continue;
                }
                if (c_ < e_ && *c_ == '}') {
                    ++c_;
                    return v;
                }
                Error("expected ',' or '}'");
            }
        }
        if (*c_ == '[') {
            ++c_;
            v.k = J::Arr;
            Skip();
            if (c_ < e_ && *c_ == ']') {
                ++c_;
                return v;
            }
            for (;;) {
                v.a.push_back(Value());
                Skip();
                if (c_ < e_ && *c_ == ',') {
                    ++c_;
                    continue;
                }
                if (c_ < e_ && *c_ == ']') {
                    ++c_;
                    return v;
                }
                Error("expected ',' or ']'");
            }
        }
        if (*c_ == '"') return J::str(String());
        if (Lit("true")) return J::boolean(true);
        if (Lit("false")) return J::boolean(false);
        if (Lit("null")) return J();
        double d = 0;
        auto r = std::from_chars(c_, e_, d);
        if (r.ec != std::errc() || r.ptr == c_) Error("unexpected character");
        c_ = r.ptr;
        return J::num(d);
    }
    const char *c_, *e_, *b_;
    std::string name_;
};

std::string ReadFile(const fs::path &p) {
    std::ifstream in(p, std::ios::binary);
    if (!in) Fail("cannot open " + p.string());
    std::stringstream ss;
    ss << in.rdbuf();
    return ss.str();
}

// Relative "file" values (and bare rig paths) are resolved against the file that names them.
// Images for environment lights and image textures (src/core/image.cpp:31-101): PFM (linear
// float, "PF" RGB or "Pf" grey, bottom row first, negative scale = little-endian), binary PPM ("P6",
// 8-bit), PNG (reduced to 8-bit RGB as stb_image does), Radiance RGBE (.hdr) and OpenEXR (.exr,
// single-part scanline files); 8-bit values are decoded from sRGB unless the image is data
// ("linear"), float formats are linear. Rows come out top row first, linear RGB. Other formats
// (JPEG, BMP, TGA, ...) are a named error.
bool ReadImageRGB(const std::string &path, int *w, int *h, std::vector<float> *rgb, std::string *err, bool srgb8) {
    std::string ext = fs::path(path).extension().string();
    for (char &c : ext) c = char(std::tolower(static_cast<unsigned char>(c)));
    FILE *f = std::fopen(path.c_str(), "rb");
    if (!f) {
        *err = "cannot open " + path;
        return false;
    }
    bool ok = false;
    if (ext == ".pfm") {
        char hdr[3] = {0};
        float scale = 0;
        ok = std::fscanf(f, "%2s %d %d %f", hdr, w, h, &scale) == 4;
        std::fgetc(f);
        const int nc = (ok && std::strcmp(hdr, "PF") == 0) ? 3 : (ok && std::strcmp(hdr, "Pf") == 0 ? 1 : 0);
        if (!ok || nc == 0 || *w <= 0 || *h <= 0 || int64_t(*w) * *h > (int64_t(1) << 26)) {
            *err = "bad PFM header";
            ok = false;
        } else {
            std::vector<float> data(size_t(*w) * size_t(*h) * size_t(nc));
            ok = std::fread(data.data(), sizeof(float), data.size(), f) == data.size();
            if (!ok) *err = "truncated PFM";
            if (ok && scale > 0)  // big-endian
                for (float &v : data) {
                    uint32_t u;
                    std::memcpy(&u, &v, 4);
                    u = (u >> 24) | ((u >> 8) & 0xff00) | ((u << 8) & 0xff0000) | (u << 24);
                    std::memcpy(&v, &u, 4);
                }
            if (ok) {
                rgb->assign(size_t(*w) * size_t(*h) * 3, 0.f);
                for (int y = 0; y < *h; ++y)
                    for (int x = 0; x < *w; ++x) {
                        const float *q = &data[(size_t(*h - 1 - y) * size_t(*w) + size_t(x)) * size_t(nc)];
                        float *o = &(*rgb)[3 * (size_t(y) * size_t(*w) + size_t(x))];
                        o[0] = q[0];
                        o[1] = nc == 3 ? q[1] : q[0];
                        o[2] = nc == 3 ? q[2] : q[0];
                    }
            }
        }
    } else if (ext == ".ppm") {
        char magic[3] = {0};
        int maxv = 0;
        ok = std::fscanf(f, "%2s", magic) == 1 && std::strcmp(magic, "P6") == 0;
        auto num = [&](int *v) {  // skips whitespace and # comments
            int c = std::fgetc(f);
            while (c != EOF && (std::isspace(c) || c == '#')) {
                if (c == '#')
                    while (c != EOF && c != '\n') c = std::fgetc(f);
                c = std::fgetc(f);
            }
            if (c == EOF || !std::isdigit(c)) return false;
            *v = 0;
            while (c != EOF && std::isdigit(c)) {
                *v = *v * 10 + (c - '0');
                c = std::fgetc(f);
            }
            return true;
        };
        ok = ok && num(w) && num(h) && num(&maxv) && maxv == 255 && *w > 0 && *h > 0 &&
             int64_t(*w) * *h <= (int64_t(1) << 26);
        if (!ok) {
            *err = "bad PPM header (binary P6, 8 bits)";
        } else {
            std::vector<unsigned char> data(size_t(*w) * size_t(*h) * 3);
            ok = std::fread(data.data(), 1, data.size(), f) == data.size();
            if (!ok) *err = "truncated PPM";
            float lut[256];
            for (int i = 0; i < 256; ++i) {  // SRGBDecode (src/core/color.h), or linear (image.cpp:96)
                float v = float(i) / 255.f;
                lut[i] = !srgb8 ? v : (v <= 0.04045f ? v / 12.92f : std::pow((v + 0.055f) / 1.055f, 2.4f));
            }
            if (ok) {
                rgb->resize(data.size());
                for (size_t i = 0; i < data.size(); ++i) (*rgb)[i] = lut[data[i]];
            }
        }
    } else if (ext == ".png" || ext == ".hdr" || ext == ".exr") {
        std::vector<unsigned char> file;
        unsigned char buf[1 << 16];
        for (size_t n; (n = std::fread(buf, 1, sizeof(buf), f)) > 0;) file.insert(file.end(), buf, buf + n);
        if (ext != ".png") {  // float formats: linear whatever the colorspace says (image.cpp:76-89)
            ok = ext == ".hdr" ? DecodeHDR(file, w, h, rgb, err) : DecodeEXR(file, w, h, rgb, err);
            std::fclose(f);
            return ok;
        }
        std::vector<unsigned char> data;
        ok = DecodePNG(file, w, h, &data, err);
        if (ok) {
            float lut[256];  // as for PPM: 8-bit values, sRGB-decoded or linear (image.cpp:96)
            for (int i = 0; i < 256; ++i) {
                float v = float(i) / 255.f;
                lut[i] = !srgb8 ? v : (v <= 0.04045f ? v / 12.92f : std::pow((v + 0.055f) / 1.055f, 2.4f));
            }
            rgb->resize(data.size());
            for (size_t i = 0; i < data.size(); ++i) (*rgb)[i] = lut[data[i]];
        }
    } else {
        *err = "image format '" + ext + "' is not read by this loader (PFM, PPM, PNG, HDR and EXR are)";
    }
    std::fclose(f);
    return ok;
}

void ResolvePaths(J &j, const fs::path &dir) {
    if (j.isObj()) {
        for (auto &[key, v] : j.o) {
            if ((key == "file" || key == "ies") && v.isStr()) {
                fs::path p(v.s);
                if (p.is_relative()) v.s = (dir / p).lexically_normal().string();
            } else if (key == "rigs" && v.isObj()) {
                for (auto &[rk, rv] : v.o) {
                    (void)rk;
                    if (rv.isStr()) {
                        fs::path p(rv.s);
                        if (p.is_relative()) rv.s = (dir / p).lexically_normal().string();
                    } else {
                        ResolvePaths(rv, dir);
                    }
                }
            } else {
                ResolvePaths(v, dir);
            }
        }
    } else if (j.isArr()) {
        for (auto &e : j.a) ResolvePaths(e, dir);
    }
}

J ReadJson(const fs::path &p) {
    J j = JsonReader(ReadFile(p), p.string()).Document();
    ResolvePaths(j, p.parent_path());
    return j;
}

const std::set<std::string> kDict = {"textures", "materials", "media", "geometry", "lights", "cameras", "rigs"};

void DeepMerge(J &dst, const J &src) {
    for (const auto &[k, v] : src.o) {
        if (dst.has(k) && dst.o[k].isObj() && v.isObj()) DeepMerge(dst.o[k], v);
        else dst.set(k) = v;
    }
}

// Dictionary sections merge by id, objects append, render settings deep-merge, the rest replaces.
void MergeDocument(J &dst, const J &src) {
    if (!dst.isObj()) dst = J::object();
    for (const auto &[k, v] : src.o) {
        if (k == "include") continue;
        if (kDict.count(k)) {
            if (!v.isObj()) Fail("section '" + k + "' must be an object");
            J &d = dst.set(k);
            if (!d.isObj()) d = J::object();
            for (const auto &[id, e] : v.o) d.o[id] = e;
        } else if (k == "objects") {
            if (!v.isArr()) Fail("'objects' must be an array");
            J &d = dst.set(k);
            if (!d.isArr()) d = J::array();
            for (const auto &e : v.a) d.a.push_back(e);
        } else if (k == "render" && dst.has("render") && v.isObj()) {
            DeepMerge(dst.o["render"], v);
        } else {
            dst.set(k) = v;
        }
    }
}

J LoadComposed(const fs::path &path, int depth = 0) {
    if (depth > 16) Fail("include depth exceeded at " + path.string());
    std::string ext = path.extension().string();
    for (char &c : ext) c = char(std::tolower((unsigned char)c));
    if (ext == ".gltf" || ext == ".glb" || ext == ".pbrt") Fail(path.string() + ": glTF and pbrt imports are not supported");
    J doc = ReadJson(path);
    J result = J::object();
    if (doc.has("include"))
        for (const auto &inc : doc.at("include").a) {
            fs::path p(inc.s);
            if (p.is_relative()) p = path.parent_path() / p;
            MergeDocument(result, LoadComposed(p, depth + 1));
        }
    if (doc.has("import")) Fail(path.string() + ": 'import' (glTF, pbrt and placed scene files) is not supported");
    MergeDocument(result, doc);
    return result;
}

float Num(const J &j, const char *what) {
    if (!j.isNum()) Fail(std::string("'") + what + "' must be a number");
    return float(j.n);
}
float GetF(const J &j, const char *key, float def) { return j.has(key) ? Num(j.at(key), key) : def; }
std::string GetS(const J &j, const char *key, const std::string &def) {
    if (!j.has(key)) return def;
    if (!j.at(key).isStr()) Fail(std::string("'") + key + "' must be a string");
    return j.at(key).s;
}
bool GetB(const J &j, const char *key, bool def) {
    if (!j.has(key)) return def;
    if (j.at(key).k != J::Bool) Fail(std::string("'") + key + "' must be true/false");
    return j.at(key).b;
}
F3 Vec3(const J &j, const char *what) {
    if (j.isNum()) return f3(float(j.n), float(j.n), float(j.n));
    if (!j.isArr() || j.size() != 3 || !j.a[0].isNum() || !j.a[1].isNum() || !j.a[2].isNum())
        Fail(std::string(what) + ": expected [x, y, z]");
    return f3(float(j.a[0].n), float(j.a[1].n), float(j.a[2].n));
}

// ---- transforms (src/core/math.h:250-315, src/core/math.cpp:71-105) -------------------------------
struct M4 {
    float m[4][4] = {{1, 0, 0, 0}, {0, 1, 0, 0}, {0, 0, 1, 0}, {0, 0, 0, 1}};
    M4 operator*(const M4 &b) const {
        M4 r;
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j) {
                float s = 0;
                for (int k = 0; k < 4; ++k) s += m[i][k] * b.m[k][j];
                r.m[i][j] = s;
            }
        return r;
    }
    M4 Transposed() const {
        M4 r;
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j) r.m[i][j] = m[j][i];
        return r;
    }
    M4 Inverse() const {  // Gauss-Jordan with partial pivoting in double
        double a[4][8];
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 8; ++j) a[i][j] = j < 4 ? m[i][j] : (j - 4 == i ? 1.0 : 0.0);
        for (int c = 0; c < 4; ++c) {
            int p = c;
            for (int r = c + 1; r < 4; ++r)
                if (std::fabs(a[r][c]) > std::fabs(a[p][c])) p = r;
            if (a[p][c] == 0) Fail("singular transform matrix");
            for (int j = 0; j < 8; ++j) std::swap(a[c][j], a[p][j]);
            double inv = 1 / a[c][c];
            for (int j = 0; j < 8; ++j) a[c][j] *= inv;
            for (int r = 0; r < 4; ++r)
                if (r != c) {
                    double f = a[r][c];
                    for (int j = 0; j < 8; ++j) a[r][j] -= f * a[c][j];
                }
        }
        M4 r;
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j) r.m[i][j] = float(a[i][j + 4]);
        return r;
    }
};

struct Xf {
    M4 m, inv;
    static Xf FromMatrix(const M4 &mat) { return {mat, mat.Inverse()}; }
    static Xf Translate(F3 t) {
        Xf x;
        x.m.m[0][3] = t.x, x.m.m[1][3] = t.y, x.m.m[2][3] = t.z;
        x.inv.m[0][3] = -t.x, x.inv.m[1][3] = -t.y, x.inv.m[2][3] = -t.z;
        return x;
    }
    static Xf Scale(F3 s) {
        Xf x;
        x.m.m[0][0] = s.x, x.m.m[1][1] = s.y, x.m.m[2][2] = s.z;
        x.inv.m[0][0] = 1 / s.x, x.inv.m[1][1] = 1 / s.y, x.inv.m[2][2] = 1 / s.z;
        return x;
    }
    // Rotation by `deg` degrees about `axis` (right-handed), a rotation matrix and its transpose.
    static Xf Rotate(float deg, F3 axis) {
        F3 a = norm3(axis);
        float rad = deg * (3.14159265358979323846f / 180.f);
        float s = std::sin(rad), c = std::cos(rad);
        M4 r;
        r.m[0][0] = a.x * a.x + (1 - a.x * a.x) * c;
        r.m[0][1] = a.x * a.y * (1 - c) - a.z * s;
        r.m[0][2] = a.x * a.z * (1 - c) + a.y * s;
        r.m[1][0] = a.x * a.y * (1 - c) + a.z * s;
        r.m[1][1] = a.y * a.y + (1 - a.y * a.y) * c;
        r.m[1][2] = a.y * a.z * (1 - c) - a.x * s;
        r.m[2][0] = a.x * a.z * (1 - c) - a.y * s;
        r.m[2][1] = a.y * a.z * (1 - c) + a.x * s;
        r.m[2][2] = a.z * a.z + (1 - a.z * a.z) * c;
        return {r, r.Transposed()};
    }
    Xf operator*(const Xf &t) const { return {m * t.m, t.inv * inv}; }
    F3 Point(F3 p) const {
        const auto &a = m.m;
        float x = a[0][0] * p.x + a[0][1] * p.y + a[0][2] * p.z + a[0][3];
        float y = a[1][0] * p.x + a[1][1] * p.y + a[1][2] * p.z + a[1][3];
        float z = a[2][0] * p.x + a[2][1] * p.y + a[2][2] * p.z + a[2][3];
        float w = a[3][0] * p.x + a[3][1] * p.y + a[3][2] * p.z + a[3][3];
        return w == 1 ? f3(x, y, z) : f3(x / w, y / w, z / w);
    }
    F3 Vector(F3 v) const {
        const auto &a = m.m;
        return f3(a[0][0] * v.x + a[0][1] * v.y + a[0][2] * v.z, a[1][0] * v.x + a[1][1] * v.y + a[1][2] * v.z,
                  a[2][0] * v.x + a[2][1] * v.y + a[2][2] * v.z);
    }
    F3 Normal(F3 n) const {  // inverse transpose
        const auto &a = inv.m;
        return f3(a[0][0] * n.x + a[1][0] * n.y + a[2][0] * n.z, a[0][1] * n.x + a[1][1] * n.y + a[2][1] * n.z,
                  a[0][2] * n.x + a[1][2] * n.y + a[2][2] * n.z);
    }
};

M4 Matrix16(const J &m, const std::string &what) {
    if (!m.isArr() || m.size() != 16) Fail(what + " must have 16 numbers (row-major)");
    M4 r;
    for (int i = 0; i < 16; ++i) r.m[i / 4][i % 4] = Num(m.a[size_t(i)], "matrix");
    return r;
}

// translate * rotate * scale, or an explicit row-major matrix (scene_loader.cpp:260-290).
Xf ParseTransform(const J &j) {
    if (!j.isObj()) return Xf();
    if (j.has("matrix")) return Xf::FromMatrix(Matrix16(j.at("matrix"), "transform.matrix"));
    Xf T, R, S;
    if (j.has("translate")) T = Xf::Translate(Vec3(j.at("translate"), "translate"));
    if (j.has("rotate")) {
        auto one = [](const J &e) -> Xf {
            if (e.isArr() && e.size() == 4)
                return Xf::Rotate(Num(e.a[0], "rotate"), f3(Num(e.a[1], "rotate"), Num(e.a[2], "rotate"), Num(e.a[3], "rotate")));
            if (e.isObj()) return Xf::Rotate(Num(e.at("angle"), "rotate.angle"), Vec3(e.at("axis"), "rotate.axis"));
            Fail("rotate: expected [angle, ax, ay, az]");
        };
        const J &r = j.at("rotate");
        if (r.isArr() && r.size() && r.a[0].isArr()) {
            for (const auto &e : r.a) R = R * one(e);
        } else {
            R = one(r);
        }
    }
    if (j.has("euler")) {
        F3 e = Vec3(j.at("euler"), "euler");
        R = R * Xf::Rotate(e.z, f3(0, 0, 1)) * Xf::Rotate(e.y, f3(0, 1, 0)) * Xf::Rotate(e.x, f3(1, 0, 0));
    }
    if (j.has("scale")) S = Xf::Scale(Vec3(j.at("scale"), "scale"));
    return T * R * S;
}

// ---- meshes ------------------------------------------------------------------------------------
struct LMesh {
    std::vector<F3> p, n;
    std::vector<float> uv;  // 2 per vertex
    std::vector<int> idx;
    bool closed = false;
};

F3 Normalized(F3 v) {
    float l = std::sqrt(dot3(v, v));
    return l > 0 ? f3(v.x / l, v.y / l, v.z / l) : v;
}

// OBJ: v / vt / vn / f (polygons fan-triangulated; negative indices relative to the end). Each
// distinct position/uv/normal triple becomes one vertex, in order of first use (mesh_io.cpp:31-157).
LMesh LoadOBJ(const std::string &path) {
    std::string text = ReadFile(path);
    std::vector<F3> P, N;
    std::vector<float> T;
    LMesh m;
    std::map<std::tuple<int, int, int>, int> remap;
    bool anyN = false, anyT = false;
    const char *c = text.data(), *end = c + text.size();
    auto space = [&] {
        while (c < end && (*c == ' ' || *c == '\t' || *c == '\r')) ++c;
    };
    auto flt = [&]() {
        space();
        float v = 0;
        auto r = std::from_chars(c, end, v);
        if (r.ec != std::errc()) v = 0;
        c = r.ptr;
        return v;
    };
    auto integer = [&](int *v) {
        auto r = std::from_chars(c, end, *v);
        if (r.ec != std::errc()) return false;
        c = r.ptr;
        return true;
    };
    auto resolve = [](int i, size_t count) { return i < 0 ? int(count) + i : i - 1; };
    std::vector<int> face;
    while (c < end) {
        space();
        if (c + 1 < end && c[0] == 'v' && (c[1] == ' ' || c[1] == '\t')) {
            c += 1;
            float x = flt(), y = flt(), z = flt();
            P.push_back(f3(x, y, z));
        } else if (c + 2 < end && c[0] == 'v' && c[1] == 'n' && (c[2] == ' ' || c[2] == '\t')) {
            c += 2;
            float x = flt(), y = flt(), z = flt();
            N.push_back(f3(x, y, z));
        } else if (c + 2 < end && c[0] == 'v' && c[1] == 't' && (c[2] == ' ' || c[2] == '\t')) {
            c += 2;
            float u = flt(), v = flt();
            T.push_back(u);
            T.push_back(v);
        } else if (c + 1 < end && c[0] == 'f' && (c[1] == ' ' || c[1] == '\t')) {
            c += 1;
            face.clear();
            for (;;) {
                space();
                if (c >= end || *c == '\n' || *c == '#') break;
                int vi = 0, ti = 0, ni = 0;
                if (!integer(&vi)) Fail("OBJ: malformed face in " + path);
                if (c < end && *c == '/') {
                    ++c;
                    if (c < end && *c != '/') integer(&ti);
                    if (c < end && *c == '/') {
                        ++c;
                        integer(&ni);
                    }
                }
                int pv = resolve(vi, P.size());
                int pt = ti ? resolve(ti, T.size() / 2) : -1;
                int pn = ni ? resolve(ni, N.size()) : -1;
                if (pv < 0 || pv >= int(P.size())) Fail("OBJ: bad vertex index in " + path);
                auto [it, fresh] = remap.try_emplace(std::make_tuple(pv, pt, pn), int(m.p.size()));
                if (fresh) {
                    m.p.push_back(P[size_t(pv)]);
                    bool hasT = pt >= 0 && pt < int(T.size() / 2);
                    m.uv.push_back(hasT ? T[2 * size_t(pt)] : 0.f);
                    m.uv.push_back(hasT ? T[2 * size_t(pt) + 1] : 0.f);
                    m.n.push_back(pn >= 0 && pn < int(N.size()) ? Normalized(N[size_t(pn)]) : f3(0, 0, 0));
                    anyT |= pt >= 0;
                    anyN |= pn >= 0;
                }
                face.push_back(it->second);
            }
            for (size_t i = 1; i + 1 < face.size(); ++i) m.idx.insert(m.idx.end(), {face[0], face[i], face[i + 1]});
        }
        while (c < end && *c != '\n') ++c;
        if (c < end) ++c;
    }
    if (!anyT) m.uv.clear();
    if (!anyN) m.n.clear();
    for (const F3 &n : m.n)
        if (n.x == 0 && n.y == 0 && n.z == 0) {  // a vertex without a normal: no shading normals at all
            m.n.clear();
            break;
        }
    if (m.idx.empty()) Fail("OBJ: no faces in " + path);
    return m;
}

// PLY, ascii or binary_little_endian: vertex x y z [nx ny nz] [u v | s t | texture_u texture_v],
// face vertex_indices (or vertex_index) lists, fan-triangulated (mesh_io.cpp:160-300).
LMesh LoadPLY(const std::string &path) {
    std::ifstream in(path, std::ios::binary);
    if (!in) Fail("cannot open " + path);
    struct Prop {
        std::string name, type, countType;
        bool list = false;
    };
    struct Elem {
        std::string name;
        size_t count = 0;
        std::vector<Prop> props;
    };
    std::vector<Elem> elems;
    std::string line, format;
    std::getline(in, line);
    if (line.rfind("ply", 0) != 0) Fail("not a PLY file: " + path);
    while (std::getline(in, line)) {
        if (!line.empty() && line.back() == '\r') line.pop_back();
        std::istringstream ss(line);
        std::string tag;
        ss >> tag;
        if (tag == "format") ss >> format;
        else if (tag == "element") {
            Elem e;
            ss >> e.name >> e.count;
            elems.push_back(e);
        } else if (tag == "property" && !elems.empty()) {
            Prop p;
            std::string t;
            ss >> t;
            if (t == "list") {
                p.list = true;
                ss >> p.countType >> p.type >> p.name;
            } else {
                p.type = t;
                ss >> p.name;
            }
            elems.back().props.push_back(p);
        } else if (tag == "end_header") {
            break;
        }
    }
    const bool ascii = format == "ascii";
    if (!ascii && format != "binary_little_endian") Fail("unsupported PLY format '" + format + "'");
    auto size = [](const std::string &t) {
        if (t == "char" || t == "uchar" || t == "int8" || t == "uint8") return 1;
        if (t == "short" || t == "ushort" || t == "int16" || t == "uint16") return 2;
        if (t == "int" || t == "uint" || t == "float" || t == "int32" || t == "uint32" || t == "float32") return 4;
        if (t == "double" || t == "float64") return 8;
        return 0;
    };
    auto read = [&](const std::string &t) -> double {
        if (ascii) {
            double v = 0;
            in >> v;
            return v;
        }
        unsigned char b[8] = {0};
        const int sz = size(t);
        if (sz == 0) Fail("PLY: unknown property type '" + t + "' in " + path);
        in.read(reinterpret_cast<char *>(b), sz);
        if (t == "char" || t == "int8") return double(int8_t(b[0]));
        if (t == "uchar" || t == "uint8") return double(b[0]);
        if (t == "short" || t == "int16") {
            int16_t v;
            std::memcpy(&v, b, 2);
            return v;
        }
        if (t == "ushort" || t == "uint16") {
            uint16_t v;
            std::memcpy(&v, b, 2);
            return v;
        }
        if (t == "int" || t == "int32") {
            int32_t v;
            std::memcpy(&v, b, 4);
            return v;
        }
        if (t == "uint" || t == "uint32") {
            uint32_t v;
            std::memcpy(&v, b, 4);
            return v;
        }
        if (t == "float" || t == "float32") {
            float v;
            std::memcpy(&v, b, 4);
            return v;
        }
        double v;
        std::memcpy(&v, b, 8);
        return v;
    };
    LMesh m;
    for (const Elem &e : elems) {
        if (e.name == "vertex") {
            bool hasN = false, hasT = false;
            for (const Prop &p : e.props) {
                hasN |= p.name == "nx";
                hasT |= p.name == "u" || p.name == "s" || p.name == "texture_u";
            }
            m.p.assign(e.count, f3(0, 0, 0));
            if (hasN) m.n.assign(e.count, f3(0, 0, 0));
            if (hasT) m.uv.assign(2 * e.count, 0.f);
            for (size_t i = 0; i < e.count; ++i)
                for (const Prop &p : e.props) {
                    if (p.list) {
                        int n = int(read(p.countType));
                        for (int k = 0; k < n; ++k) read(p.type);
                        continue;
                    }
                    float v = float(read(p.type));
                    if (p.name == "x") m.p[i].x = v;
                    else if (p.name == "y") m.p[i].y = v;
                    else if (p.name == "z") m.p[i].z = v;
                    else if (p.name == "nx") m.n[i].x = v;
                    else if (p.name == "ny") m.n[i].y = v;
                    else if (p.name == "nz") m.n[i].z = v;
                    else if (p.name == "u" || p.name == "s" || p.name == "texture_u") m.uv[2 * i] = v;
                    else if (p.name == "v" || p.name == "t" || p.name == "texture_v") m.uv[2 * i + 1] = v;
                }
        } else if (e.name == "face") {
            std::vector<int> face;
            for (size_t i = 0; i < e.count; ++i)
                for (const Prop &p : e.props) {
                    if (!p.list) {
                        read(p.type);
                        continue;
                    }
                    int n = int(read(p.countType));
                    face.assign(size_t(std::max(n, 0)), 0);
                    for (int k = 0; k < n; ++k) face[size_t(k)] = int(read(p.type));
                    if (p.name == "vertex_indices" || p.name == "vertex_index")
                        for (int k = 1; k + 1 < n; ++k)
                            m.idx.insert(m.idx.end(), {face[0], face[size_t(k)], face[size_t(k) + 1]});
                }
        } else {
            for (size_t i = 0; i < e.count; ++i)
                for (const Prop &p : e.props) {
                    if (p.list) {
                        int n = int(read(p.countType));
                        for (int k = 0; k < n; ++k) read(p.type);
                    } else {
                        read(p.type);
                    }
                }
        }
        if (!in) Fail("PLY: truncated file " + path);
    }
    for (F3 &n : m.n) n = Normalized(n);
    for (int i : m.idx)
        if (i < 0 || i >= int(m.p.size())) Fail("PLY: bad vertex index in " + path);
    return m;
}

LMesh LoadMeshFile(const std::string &path) {
    std::string ext = fs::path(path).extension().string();
    for (char &c : ext) c = char(std::tolower((unsigned char)c));
    if (ext == ".obj") return LoadOBJ(path);
    if (ext == ".ply") return LoadPLY(path);
    Fail("unsupported mesh format '" + ext + "' (obj, ply)");
}

// Unit shapes of the scene format (src/geometry/shape.cpp:302-353): a w x d rectangle in the y = 0
// plane facing +y, a disk of `segments` facing +y, and a closed box centred at the origin.
LMesh RectMesh(float w, float d) {
    LMesh m;
    float hw = w / 2, hd = d / 2;
    m.p = {f3(-hw, 0, -hd), f3(hw, 0, -hd), f3(hw, 0, hd), f3(-hw, 0, hd)};
    m.n.assign(4, f3(0, 1, 0));
    m.uv = {0, 0, 1, 0, 1, 1, 0, 1};
    m.idx = {0, 2, 1, 0, 3, 2};
    return m;
}

LMesh DiskMesh(float radius, int segments) {
    LMesh m;
    m.p.push_back(f3(0, 0, 0));
    m.n.push_back(f3(0, 1, 0));
    m.uv.insert(m.uv.end(), {0.5f, 0.5f});
    const float twoPi = 6.28318530717958647692f;
    for (int i = 0; i < segments; ++i) {
        float a = twoPi * float(i) / float(segments);
        m.p.push_back(f3(radius * std::cos(a), 0, radius * std::sin(a)));
        m.n.push_back(f3(0, 1, 0));
        m.uv.insert(m.uv.end(), {0.5f + 0.5f * std::cos(a), 0.5f + 0.5f * std::sin(a)});
    }
    for (int i = 0; i < segments; ++i) {
        int a = 1 + i, b = 1 + (i + 1) % segments;
        m.idx.insert(m.idx.end(), {0, b, a});
    }
    return m;
}

LMesh BoxMesh(F3 size) {
    LMesh m;
    F3 h = size * 0.5f;
    struct Face {
        F3 n, u, v;
    };
    const Face faces[6] = {{f3(1, 0, 0), f3(0, 0, -1), f3(0, 1, 0)}, {f3(-1, 0, 0), f3(0, 0, 1), f3(0, 1, 0)},
                           {f3(0, 1, 0), f3(1, 0, 0), f3(0, 0, -1)}, {f3(0, -1, 0), f3(1, 0, 0), f3(0, 0, 1)},
                           {f3(0, 0, 1), f3(1, 0, 0), f3(0, 1, 0)},  {f3(0, 0, -1), f3(-1, 0, 0), f3(0, 1, 0)}};
    auto mul = [](F3 a, F3 b) { return f3(a.x * b.x, a.y * b.y, a.z * b.z); };
    for (const Face &f : faces) {
        int base = int(m.p.size());
        F3 c = mul(f.n, h), U = mul(f.u, h), V = mul(f.v, h);
        m.p.push_back(c - U - V);
        m.p.push_back(c + U - V);
        m.p.push_back(c + U + V);
        m.p.push_back(c - U + V);
        for (int i = 0; i < 4; ++i) m.n.push_back(f.n);
        m.uv.insert(m.uv.end(), {0, 0, 1, 0, 1, 1, 0, 1});
        m.idx.insert(m.idx.end(), {base, base + 1, base + 2, base, base + 2, base + 3});
    }
    m.closed = true;
    return m;
}

// World-space copy: points, normals (inverse transpose, renormalised) and, under a mirroring
// transform, reversed winding (shape.cpp:146-160).
void ApplyTransform(LMesh &m, const Xf &t) {
    for (F3 &p : m.p) p = t.Point(p);
    for (F3 &n : m.n) n = Normalized(t.Normal(n));
    const auto &a = t.m.m;
    double det = double(a[0][0]) * (double(a[1][1]) * a[2][2] - double(a[1][2]) * a[2][1]) -
                 double(a[0][1]) * (double(a[1][0]) * a[2][2] - double(a[1][2]) * a[2][0]) +
                 double(a[0][2]) * (double(a[1][0]) * a[2][1] - double(a[1][1]) * a[2][0]);
    if (det < 0)
        for (size_t i = 0; i + 2 < m.idx.size(); i += 3) std::swap(m.idx[i + 1], m.idx[i + 2]);
}

// Point-in-closed-mesh by the parity of crossings along a fixed skew direction (shape.cpp:168-189).
bool MeshContains(const SceneBuilder &sb, int first, int count, F3 pt) {
    F3 d = Normalized(f3(0.5773f, 0.5774f, 0.5771f));
    auto cross = [](F3 a, F3 b) { return f3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); };
    int crossings = 0;
    for (int i = first; i < first + count; ++i) {
        const prgpu::Primitive &p = sb.prims[size_t(i)];
        F3 p0 = f3(p.p0.x, p.p0.y, p.p0.z), p1 = f3(p.p1.x, p.p1.y, p.p1.z), p2 = f3(p.p2.x, p.p2.y, p.p2.z);
        F3 e1 = p1 - p0, e2 = p2 - p0;
        F3 pvec = cross(d, e2);
        float det = dot3(e1, pvec);
        if (det == 0) continue;
        float inv = 1 / det;
        F3 tvec = pt - p0;
        float u = dot3(tvec, pvec) * inv;
        if (u < 0 || u > 1) continue;
        F3 qvec = cross(tvec, e1);
        float v = dot3(d, qvec) * inv;
        if (v < 0 || u + v > 1) continue;
        if (dot3(e2, qvec) * inv > 0) ++crossings;
    }
    return (crossings & 1) != 0;
}

// ---- the loader ----------------------------------------------------------------------------------
enum class SpecType { Albedo, Unbounded, Illuminant };


struct Tex {  // a colour input as the exporter flattens it: spectrum, checker second colour, u scale
    int spec = -1, second = -1;
    float scale = 0;
};

struct Emission {
    int spectrum = -1;
    float scale = 1;
    bool powerMode = false;
    float power = 0;
    bool twoSided = false;
    float cosPower = 0;
};

struct Object {
    std::string name;
    int first = 0, count = 0;
    F3 lo = f3(INFINITY, INFINITY, INFINITY), hi = f3(-INFINITY, -INFINITY, -INFINITY);
    F3 probe = f3(0, 0, 0);
    bool sphere = false, closed = false;
    int interior = -1, exterior = -1;
    bool autoExterior = true;
    float Volume() const { return (hi.x - lo.x) * (hi.y - lo.y) * (hi.z - lo.z); }
    bool InBounds(F3 p) const {
        return p.x >= lo.x && p.x <= hi.x && p.y >= lo.y && p.y <= hi.y && p.z >= lo.z && p.z <= hi.z;
    }
};

class Loader {
  public:
    Loader(J doc, const std::vector<std::string> *rigs, int width, int height, SceneBuilder *sb,
           std::vector<std::string> *warnings)
        : doc_(std::move(doc)), rigs_(rigs), width_(width), height_(height), sb_(*sb), warn_(*warnings) {}

    void Load() {
        for (const char *sec : {"textures", "materials", "media", "geometry", "lights", "cameras", "rigs"})
            if (!doc_.has(sec)) doc_.set(sec) = J::object();
        if (!doc_.has("objects")) doc_.set("objects") = J::array();
        std::vector<std::string> active;
        if (rigs_) active = *rigs_;
        else if (doc_.has("active_rigs"))
            for (const auto &r : doc_.at("active_rigs").a) active.push_back(r.s);
        for (const std::string &rig : active) {
            if (!doc_.at("rigs").has(rig)) Fail("unknown light rig '" + rig + "'");
            const J rj = doc_.at("rigs").at(rig);
            J rigDoc;
            if (rj.isStr()) rigDoc = LoadComposed(rj.s);
            else if (rj.isObj() && rj.has("file")) rigDoc = LoadComposed(rj.at("file").s);
            else rigDoc = rj;
            MergeDocument(doc_, rigDoc);
        }
        ParseRender();
        int idx = 0;
        const J objects = doc_.at("objects");
        for (const J &o : objects.a) {
            if (GetB(o, "visible", true)) AddObject(o, idx);
            ++idx;
        }
        const J lights = doc_.at("lights");
        for (const auto &[id, l] : lights.o) AddLight(id, l);
        for (const auto &[obj, e] : pendingEmission_) Attach(objects_[size_t(obj)], e);
        ResolveMedia();
        ParseCamera();
        if (sb_.lights.empty()) Warn("scene has no lights; the image will be black");
    }

  private:
    void Warn(const std::string &w) {
        if (std::find(warn_.begin(), warn_.end(), w) == warn_.end()) warn_.push_back(w);
    }

    // ---- spectra (scene_loader.cpp:202-258) ----
    int Illuminant(std::function<float(float)> f) {  // tabulated every 5 nm from 360 nm
        std::vector<float> l(95), v(95);
        for (int i = 0; i < 95; ++i) {
            l[size_t(i)] = 360.f + 5.f * float(i);
            v[size_t(i)] = f(l[size_t(i)]);
        }
        return sb_.Table95(v);
    }
    int D65Scaled(float s) { return sb_.Scaled(sb_.D65Illuminant(), s); }
    int Spectrum(const J &j, SpecType type, const std::string &what) {
        if (j.isNum()) {
            float v = float(j.n);
            return type == SpecType::Illuminant ? D65Scaled(v) : sb_.Constant(v);
        }
        if (j.isArr()) {
            F3 c = Vec3(j, what.c_str());
            switch (type) {
            case SpecType::Albedo:
                return sb_.RGBAlbedo(std::clamp(c.x, 0.f, 1.f), std::clamp(c.y, 0.f, 1.f), std::clamp(c.z, 0.f, 1.f));
            case SpecType::Unbounded: return sb_.RGBUnbounded(c.x, c.y, c.z);
            case SpecType::Illuminant: {
                int u = sb_.RGBUnbounded(c.x, c.y, c.z);
                const prgpu::SpectrumEx s = sb_.spectra[size_t(u)];
                const int d65 = sb_.D65Illuminant();
                const prgpu::SpectrumEx d = sb_.spectra[size_t(d65)];
                return Illuminant([&](float l) {
                    return SpectrumValue(s, sb_.knots.data(), l) * SpectrumValue(d, sb_.knots.data(), l);
                });
            }
            }
        }
        if (j.isStr()) {
            if (j.s == "D65" || j.s == "stdillum-D65") return sb_.D65Illuminant();
            int ior = NamedIOR(j.s);
            if (ior >= 0) return ior;  // an IOR as a spectrum is not renormalised (it has no luminance)
            Fail(what + ": unknown named spectrum '" + j.s + "'");
        }
        if (j.isObj()) {
            if (j.has("blackbody") || j.has("temperature"))
                return sb_.BlackbodyUnitY(Num(j.has("blackbody") ? j.at("blackbody") : j.at("temperature"), "temperature"));
            if (j.has("sellmeier")) {
                const J &s = j.at("sellmeier");
                std::vector<float> B, C;
                if (s.isArr() && s.size() % 2 == 0) {
                    size_t n = s.size() / 2;
                    for (size_t i = 0; i < n; ++i) B.push_back(Num(s.a[i], "sellmeier"));
                    for (size_t i = 0; i < n; ++i) C.push_back(Num(s.a[n + i], "sellmeier"));
                } else if (s.isObj()) {
                    for (const J &b : s.at("B").a) B.push_back(Num(b, "sellmeier.B"));
                    for (const J &c : s.at("C").a) C.push_back(Num(c, "sellmeier.C"));
                }
                if (B.empty() || B.size() != C.size() || B.size() > 3) Fail(what + ": bad sellmeier coefficients");
                return sb_.Sellmeier(B.data(), C.data(), int(B.size()));
            }
            if (j.has("cauchy")) {
                const J &c = j.at("cauchy");
                if (!c.isArr() || c.size() != 2) Fail(what + ": cauchy expects [A, B_nm2]");
                return sb_.Cauchy(Num(c.a[0], "cauchy"), Num(c.a[1], "cauchy"));
            }
            if (j.has("samples") || j.has("file")) {
                std::vector<float> l, v;
                if (j.has("samples")) {
                    for (const J &p : j.at("samples").a) {
                        if (!p.isArr() || p.size() < 2) Fail(what + ": samples are [lambda, value] pairs");
                        l.push_back(Num(p.a[0], "samples"));
                        v.push_back(Num(p.a[1], "samples"));
                    }
                } else {
                    ReadSpectrumCSV(j.at("file").s, &l, &v);
                }
                if (l.size() < 2) Fail(what + ": a sampled spectrum needs 2 samples or more");
                int s = sb_.Piecewise(l, v);
                if (type == SpecType::Illuminant) {
                    float y = sb_.SpectrumY(s);
                    if (y > 0) return sb_.Scaled(s, 1 / y);
                }
                return s;
            }
        }
        Fail(what + ": cannot parse spectrum");
    }
    static void ReadSpectrumCSV(const std::string &file, std::vector<float> *l, std::vector<float> *v) {
        std::ifstream in(file);
        if (!in) Fail("cannot open spectrum file " + file);
        std::string line;
        while (std::getline(in, line)) {
            for (char &c : line)
                if (c == ',' || c == ';' || c == '\t') c = ' ';
            std::istringstream ss(line);
            float a, b;
            if (ss >> a >> b) l->push_back(a), v->push_back(b);
        }
    }
    int NamedIOR(const std::string &name) {
        int g = sb_.Glass(name);
        if (g >= 0) return g;
        if (name == "water") return sb_.Cauchy(1.3240f, 3089.f);
        if (name == "air") return sb_.Constant(1.000293f);
        return -1;
    }
    int IOR(const J &j, float def) {
        if (j.k == J::Null) return sb_.Constant(def);
        if (j.isStr()) {
            int s = NamedIOR(j.s);
            if (s < 0) Fail("unknown IOR '" + j.s + "'");
            return s;
        }
        return Spectrum(j, SpecType::Unbounded, "ior");
    }

    // ---- textures (scene_loader.cpp:438-553, texture.h:61-124) ----
    // An image texture's record (GetImage + ParseUVMapping, scene_loader.cpp:438-479): PFM / PPM /
    // PNG, colorspace (8-bit images decoded from sRGB unless "linear"), wrap, flip_v, uv_scale, uv_offset,
    // and the RGB multiplier; identical textures share one record.
    int ImageRecord(const J &tx, const std::string &id, bool linearDefault, const float mult[3]) {
        if (tx.has("gltf")) Fail("texture '" + id + "': glTF-embedded images are not read by this loader");
        const std::string file = GetS(tx, "file", "");
        if (file.empty()) Fail("texture '" + id + "': missing 'file'");
        const std::string cs = GetS(tx, "colorspace", linearDefault ? "linear" : "srgb");
        const bool srgb = cs == "srgb" || cs == "sRGB";
        const std::string wrapS = GetS(tx, "wrap", "repeat");
        const int wrap = wrapS == "clamp" ? 1 : (wrapS == "mirror" ? 2 : 0);
        const bool flip = GetB(tx, "flip_v", true);
        float sc[2] = {1, 1}, of[2] = {0, 0};
        if (tx.has("uv_scale")) {
            const J &s = tx.at("uv_scale");
            if (s.isNum()) sc[0] = sc[1] = float(s.n);
            else sc[0] = Num(s.a.at(0), "uv_scale"), sc[1] = Num(s.a.at(1), "uv_scale");
        }
        if (tx.has("uv_offset")) of[0] = Num(tx.at("uv_offset").a.at(0), "uv_offset"), of[1] = Num(tx.at("uv_offset").a.at(1), "uv_offset");
        char key[512];
        std::snprintf(key, sizeof key, "|%d|%d|%d|%a|%a|%a|%a|%a|%a|%a", int(srgb), wrap, int(flip), sc[0], sc[1], of[0],
                      of[1], mult[0], mult[1], mult[2]);
        const std::string k = file + key;
        auto it = imageCache_.find(k);
        if (it != imageCache_.end()) return it->second;
        int w = 0, h = 0;
        std::vector<float> img;
        std::string err;
        if (!ReadImageRGB(file, &w, &h, &img, &err, srgb)) Fail("texture '" + id + "': " + err);
        const int rec = sb_.AddImageTexture(img.data(), w, h, wrap, flip, sc, of, mult);
        imageCache_[k] = rec;
        return rec;
    }
    static void RGB3(const J &j, float out[3], const std::string &what) {
        if (j.isNum()) out[0] = out[1] = out[2] = float(j.n);
        else if (j.isArr() && j.size() == 3)
            for (int i = 0; i < 3; ++i) out[i] = Num(j.a[size_t(i)], what.c_str());
        else Fail(what + ": expected a number or [r, g, b]");
    }
    // A colour input. Image textures are read for albedo slots that take one (imageOK: a
    // material's reflectance); elsewhere, and inside checkers, they are a named error.
    Tex SpecTex(const J &v, SpecType type, const std::string &what, bool imageOK = false) {
        Tex t;
        if (v.isObj() && v.has("texture")) {
            const std::string id = v.at("texture").s;
            if (!doc_.at("textures").has(id)) Fail(what + ": unknown texture '" + id + "'");
            const J &tx = doc_.at("textures").at(id);
            const std::string tt = GetS(tx, "type", "image");
            if (tt == "image" && imageOK && type == SpecType::Albedo) {
                float mult[3] = {1, 1, 1};
                if (tx.has("multiplier")) RGB3(tx.at("multiplier"), mult, "texture '" + id + "': multiplier");
                t.spec = sb_.ImageSpectrum(ImageRecord(tx, id, false, mult));
                return t;
            }
            if (tt == "checker") {
                Tex a = SpecTex(tx.has("a") ? tx.at("a") : J::num(0.8), type, what);
                Tex b = SpecTex(tx.has("b") ? tx.at("b") : J::num(0.2), type, what);
                t.spec = a.spec;
                t.second = b.spec;
                t.scale = 1;
                if (tx.has("uv_scale")) {
                    const J &s = tx.at("uv_scale");
                    t.scale = s.isNum() ? float(s.n) : Num(s.a.at(0), "uv_scale");
                }
                return t;
            }
            if (tt == "constant") return SpecTex(tx.at("value"), type, what, imageOK);
            if (tt == "image")
                Fail("texture '" + id + "': image textures are read as a material's reflectance only (" + what + ")");
            if (tt == "noise") Fail("texture '" + id + "': noise colour textures are not supported (checker and constant are)");
            Fail("texture '" + id + "': unknown type '" + tt + "'");
        }
        t.spec = Spectrum(v, type, what);
        return t;
    }
    float FloatTex(const J &v, float def, const std::string &what) {
        if (v.isNum()) return float(v.n);
        if (v.isObj() && v.has("texture")) {
            Warn("roughness textures are approximated by a constant");
            return def;
        }
        Fail(what + ": expected a number or {\"texture\": id}");
    }
    static float RoughnessToAlpha(float r) {
        float c = std::clamp(r, 0.f, 1.f);
        return c * c;
    }

    // ---- media (scene_loader.cpp:555-627; medium.cpp:130-166) ----
    int ParseMedium(const J &j, const std::string &name) {
        const std::string type = GetS(j, "type", "homogeneous");
        if (type != "homogeneous" && type != "grid" && type != "noise")
            Fail("medium '" + name + "': unknown type '" + type + "' (homogeneous, grid, noise)");
        if (type == "noise") Fail("medium '" + name + "': noise media are not supported");
        float g = GetF(j, "g", 0.f), scale = GetF(j, "scale", 1.f);
        int sa, ss;
        if (j.has("albedo") || j.has("mfp")) {
            // multiple-scattering albedo and mean free path per channel -> single-scattering
            // coefficients (van de Hulst's similarity relation, the Christensen-Burley fit)
            F3 A = j.has("albedo") ? Vec3(j.at("albedo"), "albedo") : f3(0.8f, 0.8f, 0.8f);
            F3 mfp = j.has("mfp") ? Vec3(j.at("mfp"), "mfp") : f3(1, 1, 1);
            float sigS[3], sigA[3], Ac[3] = {A.x, A.y, A.z}, Mc[3] = {mfp.x, mfp.y, mfp.z};
            for (int c = 0; c < 3; ++c) {
                float a = std::clamp(Ac[c], 0.f, 0.999f);
                float q = 4.09712f + 4.20863f * a - std::sqrt(9.59217f + 41.6808f * a + 17.7126f * a * a);
                float alpha = std::clamp(1 - q * q, 0.f, 0.9999f);
                float sigmaT = 1 / std::max(Mc[c], 1e-6f);
                sigS[c] = alpha * sigmaT;
                sigA[c] = sigmaT - sigS[c];
            }
            ss = sb_.RGBUnbounded(sigS[0], sigS[1], sigS[2]);
            sa = sb_.RGBUnbounded(sigA[0], sigA[1], sigA[2]);
        } else {
            sa = j.has("sigma_a") ? Spectrum(j.at("sigma_a"), SpecType::Unbounded, "sigma_a") : sb_.Constant(0.f);
            ss = j.has("sigma_s") ? Spectrum(j.at("sigma_s"), SpecType::Unbounded, "sigma_s") : sb_.Constant(0.f);
        }
        if (type == "homogeneous") return sb_.AddHomogeneous(sa, ss, scale, g);
        // grid: a float32 Mitsuba .vol (version 3), channels averaged
        const std::string file = GetS(j, "file", "");
        std::string ext = fs::path(file).extension().string();
        if (ext != ".vol") Fail("medium '" + name + "': only .vol grids are supported");
        int nx, ny, nz;
        float bb[6];
        std::vector<float> data = ReadVol(file, &nx, &ny, &nz, bb);
        F3 lo = f3(bb[0], bb[1], bb[2]), hi = f3(bb[3], bb[4], bb[5]);
        if (j.has("bounds")) {
            lo = Vec3(j.at("bounds").a.at(0), "bounds");
            hi = Vec3(j.at("bounds").a.at(1), "bounds");
        }
        int m = sb_.AddGrid(sa, ss, scale * GetF(j, "density_scale", 1.f), g, nx, ny, nz, data, lo, hi);
        Xf toWorld = ParseTransform(j.at("transform"));
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 4; ++c) sb_.media[size_t(m)].worldToMedium[r * 4 + c] = toWorld.inv.m[r][c];
        return m;
    }
    static std::vector<float> ReadVol(const std::string &path, int *nx, int *ny, int *nz, float bb[6]) {
        FILE *f = std::fopen(path.c_str(), "rb");
        if (!f) Fail("cannot open " + path);
        char magic[3];
        uint8_t version = 0;
        int32_t enc = 0, ch = 0;
        bool ok = std::fread(magic, 1, 3, f) == 3 && std::memcmp(magic, "VOL", 3) == 0 && std::fread(&version, 1, 1, f) == 1 &&
                  version == 3 && std::fread(&enc, 4, 1, f) == 1 && std::fread(nx, 4, 1, f) == 1 &&
                  std::fread(ny, 4, 1, f) == 1 && std::fread(nz, 4, 1, f) == 1 && std::fread(&ch, 4, 1, f) == 1 &&
                  std::fread(bb, 4, 6, f) == 6;
        if (!ok || enc != 1 || *nx <= 0 || *ny <= 0 || *nz <= 0 || ch <= 0 || int64_t(*nx) * *ny * *nz > (int64_t(1) << 31)) {
            std::fclose(f);
            Fail(path + ": not a float32 Mitsuba .vol (version 3) file");
        }
        size_t n = size_t(*nx) * size_t(*ny) * size_t(*nz);
        std::vector<float> raw(n * size_t(ch));
        ok = std::fread(raw.data(), sizeof(float), raw.size(), f) == raw.size();
        std::fclose(f);
        if (!ok) Fail(path + ": truncated voxel data");
        std::vector<float> d(n);
        for (size_t i = 0; i < n; ++i) {
            float s = 0;
            for (int c = 0; c < ch; ++c) s += raw[i * size_t(ch) + size_t(c)];
            d[i] = s / float(ch);
        }
        return d;
    }
    int Medium(const std::string &id) {
        auto it = mediumCache_.find(id);
        if (it != mediumCache_.end()) return it->second;
        if (!doc_.at("media").has(id)) Fail("unknown medium '" + id + "'");
        int m = ParseMedium(doc_.at("media").at(id), id);
        mediumCache_[id] = m;
        return m;
    }

    // ---- materials (scene_loader.cpp:629-760; exporter gpu_path.cpp:185-280) ----
    void Conductor(const J &j, int *eta, int *k, Tex *refl, float *ax, float *ay) {
        *eta = *k = -1;
        if (j.has("metal")) {
            const std::string m = j.at("metal").s;
            *eta = sb_.Metal(m, false);
            *k = sb_.Metal(m, true);
            if (*eta < 0) Fail("unknown metal '" + m + "' (known: Au, Ag, Cu, Al)");
        } else if (j.has("eta") && j.has("k")) {
            *eta = SpecTex(j.at("eta"), SpecType::Unbounded, "eta").spec;
            *k = SpecTex(j.at("k"), SpecType::Unbounded, "k").spec;
        } else {
            J def = J::array();
            for (int i = 0; i < 3; ++i) def.a.push_back(J::num(0.9));
            *refl = SpecTex(j.has("reflectance") ? j.at("reflectance") : def, SpecType::Albedo, "reflectance", true);
        }
        J r = j.has("roughness") ? j.at("roughness") : J::num(0.0);
        float ur = FloatTex(j.has("roughness_u") ? j.at("roughness_u") : r, 0.f, "roughness_u");
        float vr = FloatTex(j.has("roughness_v") ? j.at("roughness_v") : r, ur, "roughness_v");
        *ax = RoughnessToAlpha(ur);
        *ay = RoughnessToAlpha(vr);
    }
    // Normal and bump maps (scene_loader.cpp:732-752): a normal map is an image texture read as data
    // (strength, flip_green); a bump map's height is a float texture (an image channel or the mean
    // of its channels, times its multiplier) with a scale. A constant height has no gradient.
    void ShadingMaps(const J &j, const std::string &id, int m) {
        int normal = -1, bump = -1, channel = -1;
        float strength = 1, bumpScale = 0.01f;
        bool flipGreen = false;
        const float one[3] = {1, 1, 1};
        if (j.has("normal_map")) {
            const J &nm = j.at("normal_map");
            const std::string tid = nm.isStr() ? nm.s : GetS(nm, "texture", "");
            if (!doc_.at("textures").has(tid)) Fail("material '" + id + "': unknown normal map texture '" + tid + "'");
            J t = doc_.at("textures").at(tid);
            t.set("colorspace") = J::str("linear");  // normal maps are data
            if (GetS(t, "type", "image") != "image") Fail("material '" + id + "': a normal map is an image texture");
            normal = ImageRecord(t, tid, true, one);
            if (nm.isObj()) {
                strength = GetF(nm, "strength", 1.f);
                flipGreen = GetB(nm, "flip_green", false);
            }
        }
        if (j.has("bump_map")) {
            const J &bm = j.at("bump_map");
            bumpScale = GetF(bm, "scale", 0.01f);
            std::string tid, ch = "avg";
            if (bm.has("texture")) {
                tid = bm.at("texture").s;
                ch = GetS(bm, "channel", "avg");
            } else if (bm.has("height") && bm.at("height").isObj() && bm.at("height").has("texture")) {
                tid = bm.at("height").at("texture").s;
                ch = GetS(bm.at("height"), "channel", "avg");
            } else if (!(bm.has("height") && bm.at("height").isNum())) {
                Fail("material '" + id + "': bump_map needs a texture or a numeric height");
            }
            if (!tid.empty()) {
                if (!doc_.at("textures").has(tid)) Fail("material '" + id + "': unknown bump map texture '" + tid + "'");
                const J &t = doc_.at("textures").at(tid);
                if (GetS(t, "type", "image") != "image")
                    Fail("texture '" + tid + "': bump heights are read from image textures only");
                channel = ch == "r" ? 0 : ch == "g" ? 1 : ch == "b" ? 2 : -1;
                const float mu = GetF(t, "multiplier", 1.f);
                const float mult[3] = {mu, mu, mu};
                bump = ImageRecord(t, tid, true, mult);
            }
        }
        if (normal >= 0 || bump >= 0) sb_.SetShading(m, normal, strength, flipGreen, bump, channel, bumpScale);
    }
    SceneBuilder::Coat ParseCoat(const J &j) {
        SceneBuilder::Coat c;
        c.eta = IOR(j.at("coat_ior"), 1.5f);
        c.roughness = FloatTex(j.has("coat_roughness") ? j.at("coat_roughness") : J::num(0), 0.f, "coat_roughness");
        c.thickness = FloatTex(j.has("thickness") ? j.at("thickness") : J::num(0.01), 0.01f, "thickness");
        c.albedo = j.has("albedo") ? SpecTex(j.at("albedo"), SpecType::Albedo, "albedo").spec : -1;
        c.g = GetF(j, "g", 0.f);
        return c;
    }
    void SetChecker(int mat, const Tex &t) {
        sb_.materials[size_t(mat)].albedoB = t.second;
        sb_.materials[size_t(mat)].checkerScale = t.second >= 0 ? t.scale : 0.f;
    }
    int Material(const std::string &id) {
        auto it = materialCache_.find(id);
        if (it != materialCache_.end()) return it->second;
        if (!doc_.at("materials").has(id)) Fail("unknown material '" + id + "'");
        const J &j = doc_.at("materials").at(id);
        const std::string type = GetS(j, "type", "diffuse");
        auto reflectance = [&]() {
            J r = j.has("reflectance") ? j.at("reflectance") : j.has("color") ? j.at("color") : J::num(0.5);
            return SpecTex(r, SpecType::Albedo, "reflectance", true);
        };
        int m = -1;
        if (type == "diffuse") {
            Tex r = reflectance();
            m = sb_.Diffuse(r.spec);
            SetChecker(m, r);
        } else if (type == "conductor" || type == "metal") {
            int eta, k;
            Tex refl;
            float ax, ay;
            Conductor(j, &eta, &k, &refl, &ax, &ay);
            if (eta >= 0) m = sb_.Conductor(eta, k, ax, ay);
            else {
                m = sb_.ConductorReflectance(refl.spec, ax, ay);
                SetChecker(m, refl);
            }
        } else if (type == "dielectric" || type == "glass" || type == "subsurface") {
            J r = j.has("roughness") ? j.at("roughness") : J::num(0.0);
            float ur = FloatTex(j.has("roughness_u") ? j.at("roughness_u") : r, 0.f, "roughness_u");
            float vr = FloatTex(j.has("roughness_v") ? j.at("roughness_v") : r, ur, "roughness_v");
            m = sb_.Dielectric(IOR(j.at("ior"), 1.5f), RoughnessToAlpha(ur), RoughnessToAlpha(vr), GetB(j, "thin", false));
            if (type == "subsurface") {
                J mj = j.has("medium_params") ? j.at("medium_params") : j;
                mj.set("type") = J::str("homogeneous");
                materialInterior_[id] = ParseMedium(mj, id + ".interior");
            }
        } else if (type == "coated_diffuse") {
            SceneBuilder::Coat c = ParseCoat(j);
            Tex r = reflectance();
            m = sb_.CoatedDiffuse(c, r.spec);
            SetChecker(m, r);
        } else if (type == "coated_conductor") {
            SceneBuilder::Coat c = ParseCoat(j);
            int eta, k;
            Tex refl;
            float ax, ay;
            Conductor(j, &eta, &k, &refl, &ax, &ay);
            if (eta >= 0) m = sb_.CoatedConductor(c, eta, k, ax, ay);
            else {
                m = sb_.CoatedConductorReflectance(c, refl.spec, ax, ay);
                SetChecker(m, refl);
            }
        } else if (type == "interface") {
            m = sb_.Interface();
        } else {
            Fail("material '" + id + "': unknown type '" + type + "'");
        }
        if (type != "interface") {
            if (j.has("sheen")) {
                const J &s = j.at("sheen");
                Tex col = SpecTex(s.has("color") ? s.at("color") : J::num(1.0), SpecType::Albedo, "sheen.color");
                if (col.second >= 0) Warn("checker sheen colours use their first colour on the GPU");
                float rough = FloatTex(s.has("roughness") ? s.at("roughness") : J::num(0.3), 0.3f, "sheen.roughness");
                sb_.SetSheen(m, col.spec, rough, GetF(s, "weight", 1.f));
            }
            if (j.has("normal_map") || j.has("bump_map")) ShadingMaps(j, id, m);
        }
        if (j.has("medium")) materialInterior_[id] = Medium(j.at("medium").s);
        if (j.has("emission") && j.at("emission").k != J::Null) materialEmission_[id] = ParseEmission(j.at("emission"));
        materialCache_[id] = m;
        return m;
    }

    Emission ParseEmission(const J &j) {  // scene_loader.cpp:656-671
        Emission e;
        if (j.has("temperature")) e.spectrum = sb_.BlackbodyUnitY(Num(j.at("temperature"), "temperature"));
        else if (j.has("spectrum")) e.spectrum = Spectrum(j.at("spectrum"), SpecType::Illuminant, "emission.spectrum");
        else if (j.has("color")) e.spectrum = Spectrum(j.at("color"), SpecType::Illuminant, "emission.color");
        else e.spectrum = sb_.D65Illuminant();
        e.scale = GetF(j, "luminance", GetF(j, "intensity", 1.f));
        if (j.has("power")) {
            e.powerMode = true;
            e.power = GetF(j, "power", 1.f);
        }
        e.twoSided = GetB(j, "two_sided", false);
        e.cosPower = GetF(j, "cos_power", 0.f);
        return e;
    }

    // ---- objects (scene_loader.cpp:802-924) ----
    void AddObject(const J &o, int index) {
        Object rec;
        rec.name = GetS(o, "name", "object" + std::to_string(index));
        if (!o.has("geometry")) Fail("object '" + rec.name + "': missing 'geometry'");
        const std::string gid = o.at("geometry").s;
        if (!doc_.at("geometry").has(gid)) Fail("object '" + rec.name + "': unknown geometry '" + gid + "'");
        const J g = doc_.at("geometry").at(gid);
        const Xf xf = ParseTransform(o.at("transform"));
        if (!o.has("material") || o.at("material").k == J::Null)
            Fail("object '" + rec.name + "': missing 'material' (use a material of type 'interface' for pure medium boundaries)");
        const std::string matId = o.at("material").s;
        const int mat = Material(matId);
        const bool flip = GetB(o, "flip_normals", false);
        const std::string gtype = GetS(g, "type", "mesh");
        rec.first = int(sb_.prims.size());
        if (gtype == "sphere") {
            float r = GetF(g, "radius", 1.f);
            F3 c = g.has("center") ? Vec3(g.at("center"), "center") : f3(0, 0, 0);
            F3 wc = xf.Point(c);
            auto len = [](F3 v) { return std::sqrt(dot3(v, v)); };
            float lx = len(xf.Vector(f3(1, 0, 0))), ly = len(xf.Vector(f3(0, 1, 0))), lz = len(xf.Vector(f3(0, 0, 1)));
            if (std::fabs(lx - ly) > 1e-3f * lx || std::fabs(lx - lz) > 1e-3f * lx)
                Warn("object '" + rec.name + "': non-uniform scale on a sphere; using the average scale");
            float wr = r * (lx + ly + lz) / 3;
            int p = sb_.AddSphere(wc, wr, mat);
            sb_.prims[size_t(p)].flip = flip ? 1 : 0;
            rec.probe = wc;
            rec.closed = rec.sphere = true;
            rec.lo = f3(wc.x - wr, wc.y - wr, wc.z - wr);
            rec.hi = f3(wc.x + wr, wc.y + wr, wc.z + wr);
        } else {
            LMesh mesh;
            if (gtype == "rect" || gtype == "quad") {
                float w = 1, d = 1;
                if (g.has("size")) {
                    w = Num(g.at("size").a.at(0), "size");
                    d = Num(g.at("size").a.at(1), "size");
                }
                mesh = RectMesh(w, d);
            } else if (gtype == "box" || gtype == "cube") {
                mesh = BoxMesh(g.has("size") ? Vec3(g.at("size"), "size") : f3(1, 1, 1));
            } else if (gtype == "disk") {
                mesh = DiskMesh(GetF(g, "radius", 1.f), int(GetF(g, "segments", 64)));
            } else if (gtype == "mesh" && g.has("positions")) {
                const J &P = g.at("positions");
                for (size_t i = 0; i + 2 < P.size(); i += 3)
                    mesh.p.push_back(f3(Num(P.a[i], "positions"), Num(P.a[i + 1], "positions"), Num(P.a[i + 2], "positions")));
                if (g.has("indices"))
                    for (const J &v : g.at("indices").a) mesh.idx.push_back(int(Num(v, "indices")));
                else
                    for (int i = 0; i < int(mesh.p.size()); ++i) mesh.idx.push_back(i);
                if (g.has("normals")) {
                    const J &N = g.at("normals");
                    for (size_t i = 0; i + 2 < N.size(); i += 3)
                        mesh.n.push_back(Normalized(f3(Num(N.a[i], "normals"), Num(N.a[i + 1], "normals"), Num(N.a[i + 2], "normals"))));
                }
                if (g.has("uvs"))
                    for (const J &u : g.at("uvs").a) mesh.uv.push_back(Num(u, "uvs"));
                if (mesh.n.size() != mesh.p.size()) mesh.n.clear();
                if (mesh.uv.size() != 2 * mesh.p.size()) mesh.uv.clear();
                mesh.idx.resize(mesh.idx.size() / 3 * 3);
                for (int i : mesh.idx)
                    if (i < 0 || i >= int(mesh.p.size())) Fail("geometry '" + gid + "': index out of range");
                mesh.closed = GetB(g, "closed", false);
            } else if (gtype == "mesh") {
                const std::string file = GetS(g, "file", "");
                auto mc = meshCache_.find(file);
                if (mc == meshCache_.end()) mc = meshCache_.emplace(file, LoadMeshFile(file)).first;
                mesh = mc->second;
                mesh.closed = GetB(g, "closed", false);
            } else if (gtype == "gltf") {
                Fail("geometry '" + gid + "': glTF meshes are not supported");
            } else {
                Fail("geometry '" + gid + "': unknown type '" + gtype + "'");
            }
            ApplyTransform(mesh, xf);
            if (flip) {
                for (size_t i = 0; i + 2 < mesh.idx.size(); i += 3) std::swap(mesh.idx[i + 1], mesh.idx[i + 2]);
                for (F3 &n : mesh.n) n = f3(-n.x, -n.y, -n.z);
            }
            Mesh out;
            out.p = mesh.p;
            out.n = mesh.n;
            out.uv = mesh.uv;
            out.idx = mesh.idx;
            sb_.AddMesh(out, mat, -1, -1, !mesh.n.empty(), false);
            rec.closed = mesh.closed;
            for (const F3 &p : mesh.p) {
                rec.lo = f3(std::min(rec.lo.x, p.x), std::min(rec.lo.y, p.y), std::min(rec.lo.z, p.z));
                rec.hi = f3(std::max(rec.hi.x, p.x), std::max(rec.hi.y, p.y), std::max(rec.hi.z, p.z));
            }
            rec.probe = (rec.lo + rec.hi) * 0.5f;
        }
        rec.count = int(sb_.prims.size()) - rec.first;
        if (o.has("interior")) rec.interior = o.at("interior").k == J::Null ? -1 : Medium(o.at("interior").s);
        else if (materialInterior_.count(matId)) rec.interior = materialInterior_[matId];
        if (o.has("exterior")) {
            const J &e = o.at("exterior");
            if (e.isStr() && e.s == "auto") rec.autoExterior = true;
            else {
                rec.autoExterior = false;
                rec.exterior = e.k == J::Null ? -1 : Medium(e.s);
            }
        }
        if (rec.interior >= 0 && !rec.closed)
            Warn("object '" + rec.name + "' has an interior medium but is not a closed shape");
        objects_.push_back(rec);
        if (o.has("emission")) pendingEmission_.push_back({int(objects_.size() - 1), ParseEmission(o.at("emission"))});
        else if (materialEmission_.count(matId)) pendingEmission_.push_back({int(objects_.size() - 1), materialEmission_[matId]});
    }

    float PrimArea(const prgpu::Primitive &p) const {
        if (p.type == prgpu::PrimSphere) return 4 * 3.14159265358979323846f * p.r * p.r;
        F3 a = f3(p.p1.x - p.p0.x, p.p1.y - p.p0.y, p.p1.z - p.p0.z), b = f3(p.p2.x - p.p0.x, p.p2.y - p.p0.y, p.p2.z - p.p0.z);
        F3 c = f3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
        return 0.5f * std::sqrt(dot3(c, c));
    }

    // Every primitive of an emitting object is its own area light (scene_loader.cpp:785-800).
    void Attach(const Object &rec, const Emission &e) {
        float scale = e.scale;
        if (e.powerMode) {
            float area = 0;
            for (int i = rec.first; i < rec.first + rec.count; ++i) area += PrimArea(sb_.prims[size_t(i)]);
            float lobe = 2 * 3.14159265358979323846f / (e.cosPower + 2) * (e.twoSided ? 2 : 1);
            float y = sb_.SpectrumY(e.spectrum);
            scale = (area > 0 && y > 0) ? e.power / (y * area * lobe) : 0.f;
        }
        for (int i = rec.first; i < rec.first + rec.count; ++i) {
            prgpu::LightEx l{};
            l.kind = prgpu::LightExArea;
            l.primitive = i;
            l.spectrum = e.spectrum;
            l.scale = scale;
            l.twoSided = e.twoSided ? 1 : 0;
            l.cosPower = e.cosPower;
            sb_.prims[size_t(i)].light = int(sb_.lights.size());
            sb_.lights.push_back(l);
        }
    }

    bool Contains(const Object &o, F3 p) const {
        if (!o.InBounds(p)) return false;
        if (o.sphere) {
            const prgpu::Primitive &s = sb_.prims[size_t(o.first)];
            F3 d = f3(p.x - s.p0.x, p.y - s.p0.y, p.z - s.p0.z);
            return dot3(d, d) < s.r * s.r;
        }
        return o.closed && MeshContains(sb_, o.first, o.count, p);
    }

    // The smallest closed object with an interior medium that contains p (scene_loader.cpp:1058-1075).
    int EnclosingMedium(F3 p) const {
        int best = -1;
        float bestVol = INFINITY;
        for (const Object &o : objects_) {
            if (o.interior < 0 || !o.closed || !Contains(o, p)) continue;
            if (o.Volume() < bestVol) {
                bestVol = o.Volume();
                best = o.interior;
            }
        }
        return best;
    }

    // Exterior media: the smallest larger closed object with an interior that contains this
    // object's probe point (scene_loader.cpp:1077-1108).
    void ResolveMedia() {
        for (size_t i = 0; i < objects_.size(); ++i) {
            Object &o = objects_[i];
            if (o.autoExterior) {
                int best = -1;
                float bestVol = INFINITY, mine = o.Volume();
                for (size_t k = 0; k < objects_.size(); ++k) {
                    const Object &c = objects_[k];
                    if (k == i || c.interior < 0 || !c.closed) continue;
                    float vol = c.Volume();
                    if (vol <= mine || !Contains(c, o.probe)) continue;
                    if (vol < bestVol) {
                        bestVol = vol;
                        best = c.interior;
                    }
                }
                o.exterior = best;
            }
            for (int p = o.first; p < o.first + o.count; ++p) {
                sb_.prims[size_t(p)].mediumInside = o.interior >= 0 ? o.interior : o.exterior;
                sb_.prims[size_t(p)].mediumOutside = o.exterior;
            }
        }
    }

    // ---- lights (scene_loader.cpp:926-1056) ----
    int LightSpectrum(const J &l) {
        if (l.has("temperature")) return sb_.BlackbodyUnitY(Num(l.at("temperature"), "temperature"));
        if (l.has("spectrum")) return Spectrum(l.at("spectrum"), SpecType::Illuminant, "spectrum");
        if (l.has("color")) return Spectrum(l.at("color"), SpecType::Illuminant, "color");
        return sb_.D65Illuminant();
    }
    void AddLight(const std::string &id, const J &l) {
        if (!GetB(l, "enabled", true)) return;
        const std::string type = GetS(l, "type", "");
        const float scaleMul = GetF(l, "scale", 1.f);
        if (type == "point" || type == "spot") {
            int I = LightSpectrum(l);
            F3 pos = Vec3(l.at("position"), "position");
            float scale = GetF(l, "intensity", 1.f) * scaleMul;
            if (EnclosingMedium(pos) >= 0) Warn("light '" + id + "': point lights inside media are not attenuated on the GPU");
            if (type == "point" && l.has("ies")) {
                // Goniometric light: the IES nadir points along 'direction' (default down), turned by
                // 'rotation' degrees about it (scene_loader.cpp:940-955).
                IesProfile prof;
                std::string e;
                const std::string file = GetS(l, "ies", "");
                if (!ParseIES(ReadFile(file), &prof, &e)) Fail("light '" + id + "': " + file + ": " + e);
                F3 dir = l.has("look_at")     ? Vec3(l.at("look_at"), "look_at") - pos
                         : l.has("direction") ? Vec3(l.at("direction"), "direction")
                                              : f3(0, -1, 0);
                F3 z = norm3(dir);
                float sign = std::copysign(1.f, z.z);  // Frame::FromZ (math.h:119-126, 130-135)
                float a = -1 / (sign + z.z);
                float b = z.x * z.y * a;
                F3 fx = f3(1 + sign * z.x * z.x * a, sign * b, -sign * z.x);
                F3 fy = f3(b, sign + z.y * z.y * a, -z.y);
                float rot = (3.14159265358979323846f / 180.f) * GetF(l, "rotation", 0.f);
                F3 xin = fx * std::cos(rot) + fy * std::sin(rot);
                // Frame::FromXZ: x made orthogonal to z and normalised (v / s is v * (1 / s)), y = z x x
                // in double (math.h:79-82, 95-99, 137-148)
                float dz = dot3(xin, z);
                F3 x = xin - z * dz;
                float l2 = dot3(x, x);
                F3 frame[3] = {fx, fy, z};
                if (!(l2 < 1e-12f) && std::isfinite(l2)) {
                    x = x * (1.f / std::sqrt(l2));
                    frame[0] = x;
                    frame[1] = crossd(z, x);
                }
                if (GetB(l, "photometric", false)) scale *= prof.peak;
                if (l.has("power")) scale = GetF(l, "power", 1.f) / (prof.integral * sb_.SpectrumY(I));
                sb_.AddGoniometricLight(pos, frame, prof, I, scale);
                return;
            }
            if (type == "point") {
                if (l.has("power")) scale = GetF(l, "power", 1.f) / (4 * 3.14159265358979323846f * sb_.SpectrumY(I));
                sb_.AddPointLight(pos, I, scale);
            } else {
                F3 dir = l.has("look_at") ? Vec3(l.at("look_at"), "look_at") - pos : Vec3(l.at("direction"), "direction");
                float width = GetF(l, "cone_angle", 30.f), start = GetF(l, "falloff_start", width * 0.8f);
                sb_.AddSpotLight(pos, dir, I, scale, width, start);
            }
            return;
        }
        if (type == "distant" || type == "directional") {
            F3 dir = l.has("look_at") && l.has("position") ? Vec3(l.at("look_at"), "look_at") - Vec3(l.at("position"), "position")
                     : l.has("direction")                   ? Vec3(l.at("direction"), "direction")
                                                            : f3(0, -1, 0);
            float scale = GetF(l, "irradiance", GetF(l, "intensity", 1.f)) * scaleMul;
            sb_.AddDistantLight(f3(-dir.x, -dir.y, -dir.z), LightSpectrum(l), scale);
            return;
        }
        if (type == "sun_sky" || type == "sun") {  // scene_loader.cpp:977-997
            F3 sunDir;
            if (l.has("sun_direction")) sunDir = norm3(Vec3(l.at("sun_direction"), "sun_direction"));
            else {
                const float rad = 3.14159265358979323846f / 180.f;
                float el = rad * GetF(l, "elevation", 45.f), az = rad * GetF(l, "azimuth", 0.f);
                sunDir = f3(std::cos(el) * std::sin(az), std::sin(el), std::cos(el) * std::cos(az));
            }
            float turbidity = GetF(l, "turbidity", 3.f);
            float scale = GetF(l, "intensity", 0.03f) * scaleMul;
            float radius = 0.2665f * GetF(l, "sun_size", 1.f);
            if (sunDir.y > -0.02f && GetF(l, "sun_scale", 1.f) > 0)
                sb_.AddSunLight(sunDir, radius, sb_.SunRadiance(sunDir, turbidity, radius), scale * GetF(l, "sun_scale", 1.f));
            if (type == "sun_sky") {  // MakePreethamSky at resolution x resolution / 2
                const int res = int(GetF(l, "resolution", 512));
                if (res < 2) Fail("light '" + id + "': sky resolution must be at least 2");
                std::vector<float> sky(size_t(res) * size_t(res / 2) * 3);
                PreethamSky(sunDir, turbidity, GetF(l, "ground_albedo", 0.3f), res, res / 2, sky.data());
                sb_.AddEnvironmentMap(sky.data(), res, res / 2, scale * GetF(l, "sky_scale", 1.f), 0.f);
            }
            return;
        }
        if (type == "environment" || type == "env" || type == "sky") {
            if (l.has("file")) {
                int w = 0, h = 0;
                std::vector<float> img;
                std::string err;
                if (!ReadImageRGB(GetS(l, "file", ""), &w, &h, &img, &err, true)) Fail("light '" + id + "': " + err);
                sb_.AddEnvironmentMap(img.data(), w, h, GetF(l, "intensity", 1.f) * scaleMul, GetF(l, "rotation", 0.f));
                return;
            }
            const int spec = LightSpectrum(l);
            const float scale = GetF(l, "intensity", 1.f) * scaleMul;
            for (prgpu::LightEx &e : sb_.lights)
                if (e.kind == prgpu::LightExEnvironment) {
                    // Constant environments add: one environment whose radiance is the sum, tabulated.
                    const prgpu::SpectrumEx a = sb_.spectra[size_t(e.spectrum)], b = sb_.spectra[size_t(spec)];
                    const float sa = e.scale;
                    e.spectrum = Illuminant([&](float lam) {
                        return sa * SpectrumValue(a, sb_.knots.data(), lam) + scale * SpectrumValue(b, sb_.knots.data(), lam);
                    });
                    e.scale = 1;
                    return;
                }
            sb_.AddEnvironment(spec, scale);
            return;
        }
        if (type == "rect" || type == "sphere" || type == "disk") {
            F3 pos = Vec3(l.at("position"), "position");
            J o = J::object();
            o.set("name") = J::str("light:" + id);
            J geom = J::object();
            M4 m = Xf::Translate(pos).m;
            if (type == "sphere") {
                geom.set("type") = J::str("sphere");
                geom.set("radius") = J::num(GetF(l, "radius", 0.1f));
            } else {
                F3 dir = l.has("look_at")     ? Vec3(l.at("look_at"), "look_at") - pos
                         : l.has("direction") ? Vec3(l.at("direction"), "direction")
                                              : f3(0, -1, 0);
                dir = norm3(dir);
                // Frame::FromZ(dir): local +y (the emitting side) maps to dir, with (fy, dir, fx) a rotation
                float sign = std::copysign(1.f, dir.z);
                float a = -1 / (sign + dir.z);
                float b = dir.x * dir.y * a;
                F3 fx = f3(1 + sign * dir.x * dir.x * a, sign * b, -sign * dir.x);
                F3 fy = f3(b, sign + dir.y * dir.y * a, -dir.y);
                M4 r;
                r.m[0][0] = fy.x, r.m[1][0] = fy.y, r.m[2][0] = fy.z;
                r.m[0][1] = dir.x, r.m[1][1] = dir.y, r.m[2][1] = dir.z;
                r.m[0][2] = fx.x, r.m[1][2] = fx.y, r.m[2][2] = fx.z;
                m = m * r;
                if (type == "rect") {
                    geom.set("type") = J::str("rect");
                    if (l.has("size")) geom.set("size") = l.at("size");
                    else {
                        J s = J::array();
                        s.a = {J::num(1), J::num(1)};
                        geom.set("size") = s;
                    }
                } else {
                    geom.set("type") = J::str("disk");
                    geom.set("radius") = J::num(GetF(l, "radius", 0.5f));
                }
            }
            const std::string gid = "__light_geom_" + id;
            doc_.o["geometry"].o[gid] = geom;
            if (!materialCache_.count("__light_black")) materialCache_["__light_black"] = sb_.Diffuse(sb_.Constant(0.f));
            o.set("geometry") = J::str(gid);
            o.set("material") = J::str("__light_black");
            J e = J::object();
            for (const char *k : {"temperature", "spectrum", "color"})
                if (l.has(k)) e.set(k) = l.at(k);
            if (l.has("power")) e.set("power") = J::num(GetF(l, "power", 1.f) * scaleMul);
            e.set("luminance") = J::num(GetF(l, "luminance", GetF(l, "intensity", 1.f)) * scaleMul);
            e.set("two_sided") = J::boolean(GetB(l, "two_sided", false));
            e.set("cos_power") = J::num(GetF(l, "cos_power", 0.f));
            o.set("emission") = e;
            J mat = J::array();
            for (int i = 0; i < 4; ++i)
                for (int k = 0; k < 4; ++k) mat.a.push_back(J::num(m.m[i][k]));
            J t = J::object();
            t.set("matrix") = mat;
            o.set("transform") = t;
            AddObject(o, int(objects_.size()));
            return;
        }
        Fail("light '" + id + "': unknown type '" + type + "'");
    }

    // ---- render settings and camera (scene_loader.cpp:1129-1262) ----
    void ParseRender() {
        const J r = doc_.has("render") ? doc_.at("render") : J::object();
        cameraName_ = GetS(r, "camera", "");
        sb_.seed = r.has("seed") ? uint64_t(Num(r.at("seed"), "seed")) : 0;
        width_ = width_ > 0 ? width_ : 0;
        int w = 640, h = 360;
        prgpu::FilterType ft = prgpu::FilterBlackmanHarris;
        float radius = 1.5f;
        if (r.has("film")) {
            const J &f = r.at("film");
            w = int(GetF(f, "width", 640));
            h = int(GetF(f, "height", 360));
            clamp_ = std::max(0.f, GetF(f, "clamp", 0.f));
            if (f.has("filter")) {
                const J &fl = f.at("filter");
                const std::string t = fl.isStr() ? fl.s : GetS(fl, "type", "blackman_harris");
                if (t == "box") ft = prgpu::FilterBox, radius = 0.5f;
                else if (t == "gaussian") ft = prgpu::FilterGaussian;
                else if (t == "blackman_harris") ft = prgpu::FilterBlackmanHarris;
                else Fail("unknown filter '" + t + "'");
                if (fl.isObj() && fl.has("radius")) radius = Num(fl.at("radius"), "radius");
            }
        }
        width_ = width_ > 0 ? width_ : w;
        height_ = height_ > 0 ? height_ : h;
        sb_.camera.filterType = int(ft);
        sb_.camera.filterRadius = radius;
        sb_.spp = 64;
        sb_.maxDepth = -1;
        sb_.rrStartDepth = 1;
        if (r.has("integrator")) {
            const J &i = r.at("integrator");
            const std::string t = GetS(i, "type", "mmlt");
            if (t != "path" && t != "pt")
                Warn("integrator '" + t + "' is rendered with the GPU path tracer (the only integrator it has)");
            sb_.maxDepth = int(GetF(i, "max_depth", float(sb_.maxDepth)));
            sb_.spp = int(GetF(i, "spp", float(sb_.spp)));
            russianRoulette_ = GetB(i, "russian_roulette", true);
            if (i.has("light_sampler")) {
                const std::string ls = GetS(i, "light_sampler", "bvh");
                if (ls != "bvh" && ls != "power") Fail("light_sampler must be 'bvh' or 'power'");
                lightBVH_ = ls == "bvh";
            }
            if (i.has("time_limit")) {
                const J &tl = i.at("time_limit");
                if (tl.isNum()) timeLimit_ = tl.n;
                else Fail("time_limit: only a number of seconds is supported");
            }
        }
    }

    void ParseCamera() {
        const J &cams = doc_.at("cameras");
        std::string name = cameraName_;
        if (name.empty() && cams.size()) name = cams.o.begin()->first;
        if (name.empty() || !cams.has(name)) Fail("camera '" + name + "' not found");
        const J &c = cams.at(name);
        const std::string type = GetS(c, "type", "perspective");
        if (type != "perspective" && type != "thin_lens" && type != "pinhole") {
            if (type == "orthographic" || type == "ortho" || type == "realistic")
                Fail("camera '" + name + "': the GPU supports perspective/thin-lens cameras only");
            Fail("camera '" + name + "': unknown type '" + type + "'");
        }
        F3 pos, target;
        float toWorld[12];
        if (c.has("matrix")) {
            M4 m = Matrix16(c.at("matrix"), "camera '" + name + "': matrix");
            for (int r = 0; r < 3; ++r)
                for (int k = 0; k < 4; ++k) toWorld[r * 4 + k] = m.m[r][k];
            Xf x = Xf::FromMatrix(m);
            pos = x.Point(f3(0, 0, 0));
            target = pos + x.Vector(f3(0, 0, 1));
        } else {
            pos = Vec3(c.at("position"), "camera.position");
            target = c.has("look_at") ? Vec3(c.at("look_at"), "camera.look_at") : pos + Vec3(c.at("direction"), "camera.direction");
            F3 up = c.has("up") ? Vec3(c.at("up"), "camera.up") : f3(0, 1, 0);
            LookAt(pos, target, up, toWorld);
        }
        const float sensorH = GetF(c, "sensor_height_mm", 24.f);
        float fovY = 40;
        if (c.has("fov_y")) fovY = GetF(c, "fov_y", 40.f);
        else if (c.has("focal_length_mm"))
            fovY = 2 * std::atan(sensorH / (2 * GetF(c, "focal_length_mm", 50.f))) * (180.f / 3.14159265358979323846f);
        auto dist = [](F3 a, F3 b) {
            F3 d = a - b;
            return std::sqrt(dot3(d, d));
        };
        float focal = c.has("focus_on") ? dist(pos, Vec3(c.at("focus_on"), "focus_on")) : GetF(c, "focus_distance", dist(pos, target));
        float lens = 0;
        if (c.has("aperture_radius")) lens = GetF(c, "aperture_radius", 0.f);
        else if (c.has("f_stop") && type != "pinhole") {
            float focalMm = sensorH / (2 * std::tan((3.14159265358979323846f / 180.f) * fovY / 2));
            lens = (focalMm / 1000.f) / (2 * GetF(c, "f_stop", 8.f)) * GetF(c, "units_per_meter", 1.f);
        }
        if (type == "pinhole") lens = 0;
        std::memcpy(sb_.camera.cameraToWorld, toWorld, sizeof(toWorld));
        sb_.camera.tanHalfFov = std::tan(((3.14159265358979323846f / 180.f) * fovY) / 2);
        sb_.camera.lensRadius = lens;
        sb_.camera.focalDistance = focal;
        sb_.camera.width = width_;
        sb_.camera.height = height_;
        sb_.cameraMedium = EnclosingMedium(pos);
    }

  public:
    float clamp_ = 0;
    double timeLimit_ = 0;
    bool russianRoulette_ = true, lightBVH_ = true;

  private:
    J doc_;
    const std::vector<std::string> *rigs_;
    int width_, height_;
    SceneBuilder &sb_;
    std::vector<std::string> &warn_;
    std::string cameraName_;
    std::map<std::string, int> mediumCache_, materialCache_, materialInterior_;
    std::map<std::string, int> imageCache_;  // image texture records in knots, by file and parameters
    std::map<std::string, Emission> materialEmission_;
    std::map<std::string, LMesh> meshCache_;
    std::vector<Object> objects_;
    std::vector<std::pair<int, Emission>> pendingEmission_;
};

}  // namespace

bool LoadSceneFile(const std::string &path, const std::vector<std::string> *rigs, int width, int height,
                   SceneBuilder *sb, SceneFileSettings *settings, std::vector<std::string> *warnings, std::string *err) {
    try {
        Loader l(LoadComposed(fs::path(path)), rigs, width, height, sb, warnings);
        l.Load();
        if (settings) {
            settings->clampLuminance = l.clamp_;
            settings->timeLimit = l.timeLimit_;
            settings->russianRoulette = l.russianRoulette_;
            settings->lightBVH = l.lightBVH_;
        }
        return true;
    } catch (const std::exception &e) {
        *err = e.what();
        return false;
    }
}

}  // namespace phost
