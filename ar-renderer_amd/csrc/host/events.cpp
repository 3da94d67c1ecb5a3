// Machine-readable events of the render CLI (`--progress json`, `--list-devices`): one JSON object
// per line, in the form the reference CLI writes them (src/cli/main.cpp:299-301, 453-468,
// 576-587). The reference builds them with nlohmann::json and dump(): object keys in sorted order,
// no whitespace, strings in UTF-8 with control characters escaped, non-finite numbers as null and
// finite doubles as the shortest round-trip digits, printed in fixed notation for decimal exponents
// in (-4, 15] (with ".0" when integral) and as d.ddde+XX otherwise. A host in any language links
// these entry points and emits byte-identical lines.
#include "prender_host.h"

#include <dlfcn.h>

#include <charconv>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <thread>
#include <vector>

namespace {

// A JSON value restricted to what the events hold.
struct Value {
    enum Kind { Null, Bool, Int, Num, Str, Arr, Obj } kind = Null;
    bool b = false;
    long long i = 0;
    double d = 0;
    std::string s;
    std::vector<Value> arr;
    std::map<std::string, Value> obj;  // std::map keeps keys sorted, as dump() writes them
};

Value Num(double d) {
    Value v;
    v.kind = Value::Num;
    v.d = d;
    return v;
}
Value Int(long long i) {
    Value v;
    v.kind = Value::Int;
    v.i = i;
    return v;
}
Value Str(const std::string &s) {
    Value v;
    v.kind = Value::Str;
    v.s = s;
    return v;
}
Value Bool(bool b) {
    Value v;
    v.kind = Value::Bool;
    v.b = b;
    return v;
}

void PutString(std::string &o, const std::string &s) {
    o += '"';
    for (unsigned char c : s) {
        switch (c) {
        case '"': o += "\\\""; break;
        case '\\': o += "\\\\"; break;
        case '\b': o += "\\b"; break;
        case '\f': o += "\\f"; break;
        case '\n': o += "\\n"; break;
        case '\r': o += "\\r"; break;
        case '\t': o += "\\t"; break;
        default:
            if (c < 0x20) {
                char u[8];
                std::snprintf(u, sizeof(u), "\\u%04x", unsigned(c));
                o += u;
            } else {
                o += char(c);
            }
        }
    }
    o += '"';
}

// Shortest round-trip digits of d, laid out by the decimal point's position n (value = 0.digits * 10^n).
void PutDouble(std::string &o, double d) {
    if (!std::isfinite(d)) {
        o += "null";
        return;
    }
    if (d == 0) {
        o += std::signbit(d) ? "-0.0" : "0.0";
        return;
    }
    char buf[64];
    auto r = std::to_chars(buf, buf + sizeof(buf), d, std::chars_format::scientific);
    std::string sci(buf, r.ptr);  // [-]d[.ddd]e(+|-)XX
    if (sci[0] == '-') {
        o += '-';
        sci.erase(0, 1);
    }
    const size_t e = sci.find('e');
    std::string digits;
    for (size_t i = 0; i < e; ++i)
        if (sci[i] != '.') digits += sci[i];
    const int k = int(digits.size());
    const int n = std::stoi(sci.substr(e + 1)) + 1;
    constexpr int kMinExp = -4, kMaxExp = 15;
    if (k <= n && n <= kMaxExp) {
        o += digits + std::string(size_t(n - k), '0') + ".0";
    } else if (0 < n && n <= kMaxExp) {
        o += digits.substr(0, size_t(n)) + "." + digits.substr(size_t(n));
    } else if (kMinExp < n && n <= 0) {
        o += "0." + std::string(size_t(-n), '0') + digits;
    } else {
        o += digits.substr(0, 1);
        if (k > 1) o += "." + digits.substr(1);
        const int x = n - 1;
        char ex[16];
        std::snprintf(ex, sizeof(ex), "e%c%02d", x < 0 ? '-' : '+', x < 0 ? -x : x);
        o += ex;
    }
}

void Dump(std::string &o, const Value &v) {
    switch (v.kind) {
    case Value::Null: o += "null"; break;
    case Value::Bool: o += v.b ? "true" : "false"; break;
    case Value::Int: o += std::to_string(v.i); break;
    case Value::Num: PutDouble(o, v.d); break;
    case Value::Str: PutString(o, v.s); break;
    case Value::Arr: {
        o += '[';
        for (size_t i = 0; i < v.arr.size(); ++i) {
            if (i) o += ',';
            Dump(o, v.arr[i]);
        }
        o += ']';
        break;
    }
    case Value::Obj: {
        o += '{';
        bool first = true;
        for (const auto &kv : v.obj) {
            if (!first) o += ',';
            first = false;
            PutString(o, kv.first);
            o += ':';
            Dump(o, kv.second);
        }
        o += '}';
        break;
    }
    }
}

// Copies the line into out (NUL-terminated when it fits) and returns its length.
int Emit(const Value &v, char *out, int outLen) {
    std::string s;
    Dump(s, v);
    if (out && outLen > 0) {
        const size_t n = std::min(s.size(), size_t(outLen - 1));
        std::memcpy(out, s.data(), n);
        out[n] = 0;
    }
    return int(s.size());
}

Value Device(const std::string &id, const std::string &name, const std::string &backend, double memoryBytes,
             int computeUnits, bool gpu) {
    Value d;
    d.kind = Value::Obj;
    d.obj["id"] = Str(id);
    d.obj["name"] = Str(name);
    d.obj["backend"] = Str(backend);
    d.obj["memory_mb"] = Num(memoryBytes / 1048576.0);
    d.obj["compute_units"] = Int(computeUnits);
    d.obj["gpu"] = Bool(gpu);
    return d;
}

}  // namespace

// Reference ListRenderDevices (src/integrators/gpu_path.cpp:78-101) + the CLI's devices event
// (src/cli/main.cpp:453-465): the CPU first, then every device the plugin lists (at most 16).
PH_API int ph_list_devices_json(const char *pluginPath, int cpuThreads, char *out, int outLen) {
    Value list;
    list.kind = Value::Arr;
    const int threads = cpuThreads > 0 ? cpuThreads : std::max(1, int(std::thread::hardware_concurrency()));
    list.arr.push_back(Device("cpu", "CPU", "native", 0.0, threads, false));
    std::string status;
    char err[512] = {0};
    void *pl = pluginPath && *pluginPath ? ph_plugin_open(pluginPath, err, sizeof(err)) : nullptr;
    if (!pl) {
        status = pluginPath && *pluginPath ? err : "no GPU plugin path given";
    } else {
        status = "ok";
        prgpu::DeviceInfo devs[16];
        const int n = ph_plugin_list_devices(pl, devs, 16);
        for (int i = 0; i < n; ++i) {
            const std::string name(devs[i].name, strnlen(devs[i].name, sizeof(devs[i].name)));
            const std::string backend(devs[i].backend, strnlen(devs[i].backend, sizeof(devs[i].backend)));
            list.arr.push_back(Device("gpu:" + std::to_string(i), name, backend, double(devs[i].globalMemory),
                                      devs[i].computeUnits, devs[i].isGpu != 0));
        }
        ph_plugin_close(pl);
    }
    Value ev;
    ev.kind = Value::Obj;
    ev.obj["event"] = Str("devices");
    ev.obj["devices"] = list;
    ev.obj["gpu_status"] = Str(status);
    return Emit(ev, out, outLen);
}

// The CLI's progress event (src/cli/main.cpp:583-585); accept_rate only when >= 0.
PH_API int ph_progress_json(const char *stage, double fraction, double elapsed, double samplesPerPixel,
                            double acceptRate, char *out, int outLen) {
    Value ev;
    ev.kind = Value::Obj;
    ev.obj["event"] = Str("progress");
    ev.obj["stage"] = Str(stage ? stage : "");
    ev.obj["fraction"] = Num(fraction);
    ev.obj["elapsed"] = Num(elapsed);
    ev.obj["samples_per_pixel"] = Num(samplesPerPixel);
    if (acceptRate >= 0) ev.obj["accept_rate"] = Num(acceptRate);
    return Emit(ev, out, outLen);
}

// The CLI's stage event (src/cli/main.cpp:579-581) and error event (303-305).
PH_API int ph_stage_json(const char *name, char *out, int outLen) {
    Value ev;
    ev.kind = Value::Obj;
    ev.obj["event"] = Str("stage");
    ev.obj["name"] = Str(name ? name : "");
    return Emit(ev, out, outLen);
}

PH_API int ph_error_json(const char *message, char *out, int outLen) {
    Value ev;
    ev.kind = Value::Obj;
    ev.obj["event"] = Str("error");
    ev.obj["message"] = Str(message ? message : "");
    return Emit(ev, out, outLen);
}
