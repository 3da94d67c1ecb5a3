// IESNA LM-63 photometric profiles (type C) for goniometric lights: parsing, evaluation in the
// luminaire frame and the integral over the sphere, following the reference's IESProfile
// (src/lights/ies.cpp). The profile travels to the renderer as a record in the scene's knots
// (layout in prender_gpu.h, kIesHeader).
#include "scene_build.h"

#include <algorithm>
#include <cmath>
#include <sstream>

namespace phost {

namespace {

constexpr float kPiF = 3.14159265358979323846f;

// Largest index i in [0, size - 2] with pred(i) true, by bisection (src/core/math.h:45-54).
template <class P> int FindInterval(int size, const P &pred) {
    int first = 1, sz = size - 2;
    while (sz > 0) {
        int half = sz >> 1, middle = first + half;
        bool r = pred(middle);
        first = r ? middle + 1 : first;
        sz = r ? sz - (half + 1) : half;
    }
    return std::clamp(first - 1, 0, size - 2);
}

float Lerp(float t, float a, float b) { return (1 - t) * a + t * b; }

}  // namespace

// Keyword lines up to TILT=; an INCLUDE tilt block is skipped; then the photometric header and the
// angle and candela lists. Candela are scaled by the multiplier and ballast factor, clamped at 0 and
// normalised by their peak.
bool ParseIES(const std::string &text, IesProfile *p, std::string *err) {
    size_t pos = text.find("TILT=");
    if (pos == std::string::npos) {
        *err = "missing TILT= line (not an LM-63 file)";
        return false;
    }
    size_t eol = text.find('\n', pos);
    std::string tilt = text.substr(pos + 5, eol == std::string::npos ? std::string::npos : eol - pos - 5);
    std::istringstream in(eol == std::string::npos ? std::string() : text.substr(eol + 1));
    if (tilt.rfind("INCLUDE", 0) == 0) {
        float geometry;
        int n;
        in >> geometry >> n;
        for (int i = 0; i < 2 * n; ++i) {
            float skip;
            in >> skip;
        }
    }
    float nLamps, lumens, multiplier, type, units, w, l, h, ballast, future, watts;
    int nV, nH;
    in >> nLamps >> lumens >> multiplier >> nV >> nH >> type >> units >> w >> l >> h >> ballast >> future >> watts;
    if (!in || nV <= 0 || nH <= 0 || nV > 100000 || nH > 100000) {
        *err = "bad photometric header";
        return false;
    }
    if (int(type) != 1) {
        *err = "only type C photometry is supported";
        return false;
    }
    p->v.assign(size_t(nV), 0.f);
    p->h.assign(size_t(nH), 0.f);
    p->cd.assign(size_t(nV) * size_t(nH), 0.f);
    for (float &a : p->v) in >> a;
    for (float &a : p->h) in >> a;
    for (float &c : p->cd) in >> c;
    if (!in) {
        *err = "truncated candela data";
        return false;
    }
    p->peak = 0;
    for (float &c : p->cd) {
        c = std::max(0.f, c * multiplier * ballast);
        p->peak = std::max(p->peak, c);
    }
    if (p->peak <= 0) {
        *err = "all candela values are zero";
        return false;
    }
    for (float &c : p->cd) c /= p->peak;
    IesIntegral(p);
    return true;
}

// Relative intensity at vertical angle theta and horizontal angle phi (degrees); phi folds by the
// symmetry the last horizontal angle implies (0: axial, <= 90: quadrant, <= 180: bilateral).
float IesEvaluateAngles(const IesProfile &p, float theta, float phi) {
    const int nV = int(p.v.size()), nH = int(p.h.size());
    if (theta < p.v.front() || theta > p.v.back()) return 0;
    float hLast = p.h.back();
    phi = std::fmod(phi, 360.f);
    if (phi < 0) phi += 360;
    if (nH == 1 || hLast == 0) phi = 0;
    else if (hLast <= 90) {
        phi = std::fmod(phi, 180.f);
        if (phi > 90) phi = 180 - phi;
    } else if (hLast <= 180) {
        if (phi > 180) phi = 360 - phi;
    }
    auto vertical = [&](int hi) {
        if (nV == 1) return p.cd[size_t(hi) * size_t(nV)];
        int vi = FindInterval(nV, [&](int i) { return p.v[size_t(i)] <= theta; });
        float t = (theta - p.v[size_t(vi)]) / std::max(1e-6f, p.v[size_t(vi) + 1] - p.v[size_t(vi)]);
        return Lerp(std::clamp(t, 0.f, 1.f), p.cd[size_t(hi) * size_t(nV) + size_t(vi)],
                    p.cd[size_t(hi) * size_t(nV) + size_t(vi) + 1]);
    };
    if (nH == 1) return vertical(0);
    int hi = FindInterval(nH, [&](int i) { return p.h[size_t(i)] <= phi; });
    float t = (phi - p.h[size_t(hi)]) / std::max(1e-6f, p.h[size_t(hi) + 1] - p.h[size_t(hi)]);
    return Lerp(std::clamp(t, 0.f, 1.f), vertical(hi), vertical(hi + 1));
}

float IesEvaluate(const IesProfile &p, F3 w) {
    float theta = (180.f / kPiF) * std::acos(std::clamp(w.z, -1.f, 1.f));
    float phi = (180.f / kPiF) * std::atan2(w.y, w.x);
    return IesEvaluateAngles(p, theta, phi);
}

// Mean of the profile over a 128 x 64 grid of equal-area cells in (phi, cos theta), four
// sub-samples per cell, times 4 pi (ies.cpp:107-131).
void IesIntegral(IesProfile *p) {
    constexpr int NU = 128, NV = 64;
    double sum = 0;
    for (int v = 0; v < NV; ++v)
        for (int u = 0; u < NU; ++u) {
            float s = 0;
            for (int k = 0; k < 4; ++k) {
                float uu = (float(u) + ((k & 1) ? 0.75f : 0.25f)) / NU;
                float vv = (float(v) + ((k & 2) ? 0.75f : 0.25f)) / NV;
                float cosT = 1 - 2 * vv, phi = 2 * kPiF * uu;
                float sinT = std::sqrt(std::max(0.f, 1 - cosT * cosT));
                s += IesEvaluate(*p, f3(sinT * std::cos(phi), sinT * std::sin(phi), cosT)) / 4;
            }
            sum += s;
        }
    p->integral = float(sum / (double(NU) * NV) * 4 * kPiF);
}

int SceneBuilder::AddGoniometricLight(F3 pos, const F3 frame[3], const IesProfile &p, int spectrum, float scale) {
    prgpu::LightEx l{};
    l.kind = prgpu::LightExGoniometric;
    l.primitive = -1;
    l.spectrum = spectrum;
    l.scale = scale;
    l.position = pf(pos);
    l.direction = pf(frame[2]);
    l.profileOffset = int(knots.size());
    for (int a = 0; a < 3; ++a) knots.insert(knots.end(), {frame[a].x, frame[a].y, frame[a].z});
    knots.insert(knots.end(), {float(p.v.size()), float(p.h.size()), p.integral});
    knots.insert(knots.end(), p.v.begin(), p.v.end());
    knots.insert(knots.end(), p.h.begin(), p.h.end());
    knots.insert(knots.end(), p.cd.begin(), p.cd.end());
    l.profileSize = int(knots.size()) - l.profileOffset;
    lights.push_back(l);
    return int(lights.size() - 1);
}

}  // namespace phost
