// Launch-shape choices of the kernels, picked per size class by the project's autotuner
// (fastwarp_amd/autotune.py times the candidates of each knob on the device and writes the
// winners to fastwarp_amd/tuning/gfx950.json, which every context loads at creation). A knob
// without an entry for a size class keeps the library's default.
//
// Keys are "<knob>@<class>" (class = floor(log2(size)), size being the voxels of the grid or the
// samples of the level the launch covers) or "<knob>" for every size. The file is one flat JSON
// object of such keys with integer values; other keys (the autotuner's notes) are ignored.
#pragma once

#include <cctype>
#include <cstdint>
#include <fstream>
#include <map>
#include <sstream>
#include <string>

namespace fw {

class TuningTable {
 public:
  static int size_class(uint64_t n) {
    int c = 0;
    while (n > 1) {
      n >>= 1;
      ++c;
    }
    return c;
  }

  int64_t get(const std::string& knob, uint64_t size, int64_t fallback) const {
    auto it = values_.find(knob + "@" + std::to_string(size_class(size)));
    if (it != values_.end()) return it->second;
    it = values_.find(knob);
    return it != values_.end() ? it->second : fallback;
  }

  // value < 0 removes the entry.
  void set(const std::string& key, int64_t value) {
    if (value < 0) values_.erase(key);
    else values_[key] = value;
  }

  size_t size() const { return values_.size(); }

  // The knobs the library reads and the values each accepts.
  static bool valid(const std::string& knob, int64_t v) {
    if (knob == "cc.tile_z") return v >= 1 && v <= 256;
    if (knob == "cc.min_blocks") return v >= 1 && v <= 65536;
    if (knob == "inv.max_blocks") return v >= 1 && v <= 65535;
    if (knob == "mi.threads") return v == 0 || v == 256 || v == 512 || v == 1024;
    if (knob == "mi.max_blocks") return v >= 1 && v <= 65535;
    if (knob == "mi.batch_blocks") return v >= 1 && v <= 65535;
    if (knob == "mi.deriv_blocks") return v >= 1 && v <= 65535;
    if (knob == "gauss.one_pass") return v == 0 || v == 1;
    if (knob == "deriche.lds") return v == 0 || v == 1;
    if (knob == "mi.fused_slots") return v >= 0 && v <= 65535;
    if (knob == "batch.graphs") return v == 0 || v == 1;
    return false;
  }

  // Reads {"key": integer, ...}; entries whose value is not an integer are skipped. Returns false
  // when the file cannot be read or is not a flat JSON object.
  bool load(const std::string& path) {
    std::ifstream in(path);
    if (!in) return false;
    std::stringstream ss;
    ss << in.rdbuf();
    const std::string t = ss.str();
    size_t i = 0;
    auto ws = [&] {
      while (i < t.size() && std::isspace(static_cast<unsigned char>(t[i]))) ++i;
    };
    ws();
    if (i >= t.size() || t[i] != '{') return false;
    ++i;
    std::map<std::string, int64_t> out;
    for (;;) {
      ws();
      if (i < t.size() && t[i] == '}') break;
      if (i >= t.size() || t[i] != '"') return false;
      const size_t k0 = ++i;
      while (i < t.size() && t[i] != '"') ++i;
      if (i >= t.size()) return false;
      const std::string key = t.substr(k0, i - k0);
      ++i;
      ws();
      if (i >= t.size() || t[i] != ':') return false;
      ++i;
      ws();
      // the value: an integer, or anything else up to the next top-level comma or brace
      const size_t v0 = i;
      int depth = 0;
      bool quoted = false;
      for (; i < t.size(); ++i) {
        const char ch = t[i];
        if (quoted) {
          if (ch == '\\') ++i;
          else if (ch == '"') quoted = false;
          continue;
        }
        if (ch == '"') quoted = true;
        else if (ch == '[' || ch == '{') ++depth;
        else if (ch == ']' || ch == '}') {
          if (depth == 0) break;
          --depth;
        } else if (ch == ',' && depth == 0) break;
      }
      std::string v = t.substr(v0, i - v0);
      while (!v.empty() && std::isspace(static_cast<unsigned char>(v.back()))) v.pop_back();
      bool integer = !v.empty();
      for (size_t q = 0; q < v.size(); ++q)
        if (!(std::isdigit(static_cast<unsigned char>(v[q])) || (q == 0 && v[q] == '-'))) integer = false;
      if (integer) {
        const int64_t value = std::stoll(v);
        const std::string knob = key.substr(0, key.find('@'));
        if (valid(knob, value)) out[key] = value;
      }
      if (i < t.size() && t[i] == ',') ++i;
    }
    values_ = std::move(out);
    return true;
  }

 private:
  std::map<std::string, int64_t> values_;
};

}  // namespace fw
