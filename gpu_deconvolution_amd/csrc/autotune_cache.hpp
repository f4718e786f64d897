// Persistent record of autotuner picks: one text line per tuned
// configuration, "key v0 v1 v2 v3 v4 t0 ... t14" (chosen variant per kernel
// kind, then the mean launch time in ms of kind k, variant v at 3k + v,
// negative where that variant does not exist).  Off unless a path is set
// (fmvd_set_autotune_cache or FMVD_AUTOTUNE_CACHE); the file is read on every
// lookup and appended to after every tuning, so concurrent processes at worst
// tune the same configuration twice.
#pragma once
#include <string>

namespace fmvd {

struct TuneRecord {
    int chosen[5];
    float times[15];
};

// The cache file in effect ("" = none).
std::string autotune_cache_path();
void set_autotune_cache_path(const char* path);
// Looks up `key` (the last record for it wins); false when absent or off.
bool autotune_cache_find(const std::string& key, TuneRecord& rec);
// Appends a record (no-op when off; a write failure is not an error).
void autotune_cache_store(const std::string& key, const TuneRecord& rec);
// The record of one line, if it is well formed.
bool autotune_cache_parse(const std::string& line, std::string& key, TuneRecord& rec);
std::string autotune_cache_format(const std::string& key, const TuneRecord& rec);

}  // namespace fmvd
