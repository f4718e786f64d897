// Autotuner cache file (see autotune_cache.hpp).
#include "autotune_cache.hpp"

#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <mutex>
#include <sstream>

#include "../../include/fmvd_deconvolve.h"
#include "hip_check.hpp"

namespace fmvd {

namespace {
std::mutex g_mu;
bool g_set = false;
std::string g_path;
}  // namespace

std::string autotune_cache_path()
{
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_set) return g_path;
    const char* e = getenv("FMVD_AUTOTUNE_CACHE");
    return e ? std::string(e) : std::string();
}

void set_autotune_cache_path(const char* path)
{
    std::lock_guard<std::mutex> lk(g_mu);
    g_set = true;
    g_path = path ? path : "";
}

bool autotune_cache_parse(const std::string& line, std::string& key, TuneRecord& rec)
{
    std::istringstream in(line);
    if (!(in >> key)) return false;
    for (int& c : rec.chosen)
        if (!(in >> c) || c < 0 || c > 2) return false;
    for (float& t : rec.times)
        if (!(in >> t)) return false;
    std::string extra;
    return !(in >> extra);
}

std::string autotune_cache_format(const std::string& key, const TuneRecord& rec)
{
    std::ostringstream out;
    out << key;
    for (int c : rec.chosen) out << ' ' << c;
    out.precision(6);
    for (float t : rec.times) out << ' ' << t;
    return out.str();
}

bool autotune_cache_find(const std::string& key, TuneRecord& rec)
{
    const std::string path = autotune_cache_path();
    if (path.empty()) return false;
    std::ifstream in(path);
    bool found = false;
    std::string line, k;
    TuneRecord r;
    while (std::getline(in, line))
        if (autotune_cache_parse(line, k, r) && k == key) {
            rec = r;
            found = true;
        }
    return found;
}

void autotune_cache_store(const std::string& key, const TuneRecord& rec)
{
    const std::string path = autotune_cache_path();
    if (path.empty()) return;
    std::lock_guard<std::mutex> lk(g_mu);
    if (FILE* f = fopen(path.c_str(), "a")) {
        fprintf(f, "%s\n", autotune_cache_format(key, rec).c_str());
        fclose(f);
    }
}

}  // namespace fmvd

extern "C" {

void fmvd_set_autotune_cache(const char* path) { fmvd::set_autotune_cache_path(path); }

int fmvd_autotune_cache_selftest(void)
{
    // format -> parse round trip and malformed-line rejection (no device needed)
    fmvd::TuneRecord a{{0, 1, 2, 1, 0}, {0.5f, 0.25f, -1.f, 1.f, 2.f, 3.f, 4.f, 5.f, 6.f, 7.f, 8.f, 9.f, 10.f, -1.f, -1.f}};
    std::string k;
    fmvd::TuneRecord b;
    if (!fmvd::autotune_cache_parse(fmvd::autotune_cache_format("gfx950|256|2560x2560|u16|c1|p16", a), k, b)) return 1;
    if (k != "gfx950|256|2560x2560|u16|c1|p16") return 2;
    for (int i = 0; i < 5; ++i)
        if (a.chosen[i] != b.chosen[i]) return 3;
    for (int i = 0; i < 15; ++i)
        if (a.times[i] != b.times[i]) return 4;
    if (fmvd::autotune_cache_parse("key 0 1 2", k, b)) return 5;
    if (fmvd::autotune_cache_parse("key 0 1 2 3 0 1 1 1 1 1 1 1 1 1 1 1 1 1 1 1", k, b)) return 6;  // variant 3
    if (fmvd::autotune_cache_parse(fmvd::autotune_cache_format("k", a) + " extra", k, b)) return 7;
    return 0;
}

}  // extern "C"
