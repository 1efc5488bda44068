// The process-wide tuning table (launch shapes and tile choices read by batch_create) and its
// text file format: one "key value" pair per line, '#' starts a comment, keys not listed keep
// their defaults.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>

#include "pm3_host.hpp"

namespace pm3host {

namespace {

struct Field {
  const char* key;
  int Tuning::*member;
  int lo, hi;
};

const Field kFields[] = {
    {"pair_block", &Tuning::pair_block, 64, 256},
    {"pair_blocks_per_cu", &Tuning::pair_blocks_per_cu, 1, 64},
    {"sp2_tile16_n", &Tuning::sp2_tile16_n, 0, 1 << 20},
    {"sp2_lds", &Tuning::sp2_lds, 0, 6},
    {"sp2_tight_n", &Tuning::sp2_tight_n, 0, 1 << 20},
    {"large_syrk", &Tuning::large_syrk, 0, 1},
    {"large_split", &Tuning::large_split, 0, 4},
    {"wg128_nao", &Tuning::wg128_nao, 0, 1 << 20},
    {"wg256_nao", &Tuning::wg256_nao, 0, 1 << 20},
    {"solo_large_nao", &Tuning::solo_large_nao, 0, 1 << 20},
};

Tuning initial_table() {
  Tuning t;
  if (const char* path = std::getenv("PM3_TUNING_FILE")) {
    std::string err;
    Tuning f;
    if (tuning_load(path, f, err)) t = f;
    else std::fprintf(stderr, "pm3: PM3_TUNING_FILE ignored: %s\n", err.c_str());
  }
  return t;
}

}  // namespace

Tuning& tuning_table() {
  static Tuning t = initial_table();
  return t;
}

bool tuning_valid(const Tuning& t, std::string& why) {
  for (const Field& f : kFields) {
    const int v = t.*(f.member);
    if (v < f.lo || v > f.hi) {
      why = std::string("tuning entry ") + f.key + " = " + std::to_string(v) + " outside [" + std::to_string(f.lo) +
            ", " + std::to_string(f.hi) + "]";
      return false;
    }
  }
  if (t.pair_block % 64 != 0) {
    why = "tuning entry pair_block must be a multiple of 64";
    return false;
  }
  return true;
}

bool tuning_load(const char* path, Tuning& t, std::string& err) {
  std::ifstream in(path);
  if (!in) {
    err = std::string("cannot open ") + path;
    return false;
  }
  Tuning r;
  std::string line;
  int lineno = 0;
  while (std::getline(in, line)) {
    ++lineno;
    const size_t h = line.find('#');
    if (h != std::string::npos) line.resize(h);
    std::istringstream ls(line);
    std::string key;
    long long v;
    if (!(ls >> key)) continue;
    if (!(ls >> v)) {
      err = std::string(path) + ":" + std::to_string(lineno) + ": missing value for " + key;
      return false;
    }
    const Field* hit = nullptr;
    for (const Field& f : kFields)
      if (key == f.key) hit = &f;
    if (!hit) {
      err = std::string(path) + ":" + std::to_string(lineno) + ": unknown key " + key;
      return false;
    }
    r.*(hit->member) = (int)v;
  }
  if (!tuning_valid(r, err)) return false;
  t = r;
  return true;
}

bool tuning_save(const char* path, const Tuning& t, std::string& err) {
  std::ofstream out(path);
  if (!out) {
    err = std::string("cannot write ") + path;
    return false;
  }
  out << "# pm3 tuning table (launch shapes and tile choices; results agree whatever the values)\n";
  for (const Field& f : kFields) out << f.key << " " << t.*(f.member) << "\n";
  return (bool)out;
}

}  // namespace pm3host
