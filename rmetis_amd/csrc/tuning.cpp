// tuning.cpp -- the tuning table (tuning.hpp): compiled-in defaults, replaced
// on first use by tuning.json beside librmetis_hip.so (or the file named by
// RMETIS_HIP_TUNING) when it is present and valid.
#include "tuning.hpp"

#include <dlfcn.h>

#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

namespace rk {

namespace {

// Defaults: the autotuner's table for MI355X at the time of writing (the
// same values as the committed tuning.json).
constexpr int32_t kDefaults[kTuneCount] = {3, 3, 3, 3, 3, 3,  1, 1, 1, 1, 3, 3,  0, 0, 1, 1, 3, 3,
                                           0, 0, 1, 1, 3, 3,  1 << 17, 24};

std::atomic<int32_t> g_tab[kTuneCount];
std::once_flag g_once;
std::string g_source = "built-in defaults";

bool valid(int i, int32_t v) {
    if (i < kLevelBuckets) return v >= 0 && v < kSuitorShapes;
    return v >= 1;
}

// The integers of the JSON array that follows "key" in text (no nesting).
bool json_ints(const std::string& text, const char* key, std::vector<long long>* out) {
    const std::string k = std::string("\"") + key + "\"";
    size_t p = text.find(k);
    if (p == std::string::npos) return false;
    p = text.find(':', p + k.size());
    if (p == std::string::npos) return false;
    ++p;
    while (p < text.size() && (text[p] == ' ' || text[p] == '\n' || text[p] == '\t' || text[p] == '\r')) ++p;
    out->clear();
    if (p < text.size() && text[p] == '[') {
        const size_t e = text.find(']', p);
        if (e == std::string::npos) return false;
        const char* c = text.c_str() + p + 1;
        const char* end = text.c_str() + e;
        while (c < end) {
            char* nx = nullptr;
            const long long v = std::strtoll(c, &nx, 10);
            if (nx == c) {
                ++c;
                continue;
            }
            out->push_back(v);
            c = nx;
        }
        return true;
    }
    char* nx = nullptr;
    const long long v = std::strtoll(text.c_str() + p, &nx, 10);
    if (nx == text.c_str() + p) return false;
    out->push_back(v);
    return true;
}

std::string library_dir() {
    Dl_info info{};
    if (dladdr(reinterpret_cast<void*>(&tuning_source), &info) == 0 || !info.dli_fname) return "";
    std::string p = info.dli_fname;
    const size_t s = p.rfind('/');
    return s == std::string::npos ? std::string(".") : p.substr(0, s);
}

void load() {
    for (int i = 0; i < kTuneCount; ++i) g_tab[i].store(kDefaults[i], std::memory_order_relaxed);
    const char* env = std::getenv("RMETIS_HIP_TUNING");
    const std::string path = env && *env ? std::string(env) : library_dir() + "/tuning.json";
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) return;
    std::string text;
    char buf[4096];
    size_t got;
    while ((got = std::fread(buf, 1, sizeof(buf), f)) > 0) text.append(buf, got);
    std::fclose(f);
    int32_t v[kTuneCount];
    for (int i = 0; i < kTuneCount; ++i) v[i] = kDefaults[i];
    std::vector<long long> xs;
    bool ok = true;
    if (json_ints(text, "suitor_shape_by_level_bucket", &xs) && (int)xs.size() == kLevelBuckets)
        for (int i = 0; i < kLevelBuckets; ++i) v[i] = (int32_t)xs[i];
    else
        ok = false;
    if (json_ints(text, "incremental_min_n", &xs) && xs.size() == 1) v[kTuneIncMinN] = (int32_t)xs[0];
    else ok = false;
    if (json_ints(text, "spec_min_per_part", &xs) && xs.size() == 1) v[kTuneSpecMin] = (int32_t)xs[0];
    else ok = false;
    for (int i = 0; i < kTuneCount && ok; ++i) ok = valid(i, v[i]);
    if (!ok) {
        std::fprintf(stderr, "librmetis_hip: %s is not a valid tuning table; using the built-in defaults\n",
                     path.c_str());
        return;
    }
    for (int i = 0; i < kTuneCount; ++i) g_tab[i].store(v[i], std::memory_order_relaxed);
    g_source = path;
}

}  // namespace

int32_t tune(int index) {
    std::call_once(g_once, load);
    return g_tab[index].load(std::memory_order_relaxed);
}

bool set_tuning(const int32_t* values, int count, const char** why) {
    std::call_once(g_once, load);
    if (!values || count < 0 || count > kTuneCount) {
        *why = "tuning: values must hold 0 to 26 entries";
        return false;
    }
    for (int i = 0; i < count; ++i)
        if (!valid(i, values[i])) {
            *why = i < kLevelBuckets ? "tuning: a suitor shape must be 0..3" : "tuning: thresholds must be positive";
            return false;
        }
    for (int i = 0; i < count; ++i) g_tab[i].store(values[i], std::memory_order_relaxed);
    return true;
}

int level_bucket(int64_t n, int64_t m) {
    int sb = 0;
    while (sb < kSizeBuckets - 1 && n >= (int64_t)1 << (19 + 2 * sb)) ++sb;
    const double avg = n > 0 ? (double)m / (double)n : 0.0;
    static const double edges[kDegBuckets - 1] = {6, 10, 16, 24, 48};
    int db = 0;
    while (db < kDegBuckets - 1 && avg >= edges[db]) ++db;
    return sb * kDegBuckets + db;
}

const char* tuning_source() {
    std::call_once(g_once, load);
    return g_source.c_str();
}

}  // namespace rk
