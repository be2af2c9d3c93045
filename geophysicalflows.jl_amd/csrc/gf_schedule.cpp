// The schedule table and its JSON form:
//   {"format": "gf-schedules-1", "entries": [
//     {"model": 0, "nlayers": 1, "nx": 2048, "ny": 2048, "dtype": 0, "p": [8, 8, 8], "split": [1, 0, 1],
//      "step_us": 381.9}, ...]}
// ("split" may be absent: no splits; a 2-entry "split" leaves K3's column split off).
#include "gf_schedule.hpp"

#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <map>
#include <mutex>
#include <sstream>
#include <stdexcept>
#include <vector>

namespace gf {

namespace {
std::mutex mu;
std::map<SchedKey, Sched>& table() {
  static std::map<SchedKey, Sched> t;
  return t;
}

bool offered(int n, int p) { return p == 0 || (p == 4 && n >= 16) || (p == 8 && n >= 32) || (p == 16 && n >= 1024); }

// A small JSON reader for the table's own format: objects, arrays, numbers and strings.
struct Value {
  enum Kind { NUM, STR, ARR, OBJ } kind = NUM;
  double num = 0;
  std::string str;
  std::vector<Value> arr;
  std::vector<std::pair<std::string, Value>> obj;
  const Value* get(const std::string& k) const {
    for (auto& kv : obj) if (kv.first == k) return &kv.second;
    return nullptr;
  }
};

struct Reader {
  const std::string& s;
  size_t i = 0;
  explicit Reader(const std::string& src) : s(src) {}
  [[noreturn]] void fail(const char* what) const {
    throw std::runtime_error(std::string("schedule table: ") + what + " at byte " + std::to_string(i));
  }
  void ws() { while (i < s.size() && std::isspace((unsigned char)s[i])) ++i; }
  bool eat(char c) { ws(); if (i < s.size() && s[i] == c) { ++i; return true; } return false; }
  std::string string() {
    if (!eat('"')) fail("expected a string");
    std::string out;
    while (i < s.size() && s[i] != '"') {
      if (s[i] == '\\') fail("escapes are not used in this format");
      out += s[i++];
    }
    if (i >= s.size()) fail("unterminated string");
    ++i;
    return out;
  }
  Value value(int depth = 0) {
    if (depth > 8) fail("nesting too deep");
    ws();
    if (i >= s.size()) fail("unexpected end");
    Value v;
    if (s[i] == '{') {
      ++i;
      v.kind = Value::OBJ;
      if (eat('}')) return v;
      do {
        std::string k = string();
        if (!eat(':')) fail("expected ':'");
        v.obj.emplace_back(k, value(depth + 1));
      } while (eat(','));
      if (!eat('}')) fail("expected '}'");
    } else if (s[i] == '[') {
      ++i;
      v.kind = Value::ARR;
      if (eat(']')) return v;
      do v.arr.push_back(value(depth + 1)); while (eat(','));
      if (!eat(']')) fail("expected ']'");
    } else if (s[i] == '"') {
      v.kind = Value::STR;
      v.str = string();
    } else {
      const char* b = s.c_str() + i;
      char* e = nullptr;
      v.num = std::strtod(b, &e);
      if (e == b) fail("expected a value");
      i += (size_t)(e - b);
    }
    return v;
  }
};

int as_int(const Value* v, const char* name) {
  if (!v || v->kind != Value::NUM || v->num != (double)(int)v->num)
    throw std::runtime_error(std::string("schedule table: entry field '") + name + "' must be an integer");
  return (int)v->num;
}
}  // namespace

bool sched_lookup(const SchedKey& key, Sched& out) {
  std::lock_guard<std::mutex> g(mu);
  auto it = table().find(key);
  if (it == table().end()) return false;
  out = it->second;
  return true;
}

void sched_store(const SchedKey& key, const Sched& s) {
  std::lock_guard<std::mutex> g(mu);
  table()[key] = s;
}

void sched_clear() {
  std::lock_guard<std::mutex> g(mu);
  table().clear();
}

void sched_save(const std::string& path) {
  std::ostringstream o;
  {
    std::lock_guard<std::mutex> g(mu);
    o << "{\"format\": \"gf-schedules-1\", \"entries\": [";
    bool first = true;
    char buf[64];
    for (auto& kv : table()) {
      const SchedKey& k = kv.first;
      const Sched& e = kv.second;
      o << (first ? "\n" : ",\n") << "  {\"model\": " << k[0] << ", \"nlayers\": " << k[1] << ", \"nx\": " << k[2]
        << ", \"ny\": " << k[3] << ", \"dtype\": " << k[4] << ", \"p\": [" << e.p[0] << ", " << e.p[1] << ", "
        << e.p[2] << "], \"split\": [" << e.split[0] << ", " << e.split[1] << ", " << e.split[2]
        << "], \"step_us\": ";
      std::snprintf(buf, sizeof buf, "%.3f", e.step_us);
      o << buf << "}";
      first = false;
    }
    o << "\n]}\n";
  }
  std::ofstream f(path, std::ios::binary | std::ios::trunc);
  if (!f) throw std::runtime_error("schedule table: cannot open " + path + " for writing");
  f << o.str();
  if (!f) throw std::runtime_error("schedule table: write to " + path + " failed");
}

int sched_load(const std::string& path) {
  std::ifstream f(path, std::ios::binary);
  if (!f) throw std::runtime_error("schedule table: cannot open " + path);
  std::stringstream ss;
  ss << f.rdbuf();
  const std::string text = ss.str();
  Reader r(text);
  const Value root = r.value();
  r.ws();
  if (r.i != text.size()) r.fail("trailing characters");
  const Value* fmt = root.kind == Value::OBJ ? root.get("format") : nullptr;
  if (!fmt || fmt->kind != Value::STR || fmt->str != "gf-schedules-1")
    throw std::runtime_error("schedule table: not a gf-schedules-1 file");
  const Value* ents = root.get("entries");
  if (!ents || ents->kind != Value::ARR) throw std::runtime_error("schedule table: 'entries' must be an array");
  std::vector<std::pair<SchedKey, Sched>> got;
  for (const Value& e : ents->arr) {
    if (e.kind != Value::OBJ) throw std::runtime_error("schedule table: an entry is not an object");
    SchedKey k{as_int(e.get("model"), "model"), as_int(e.get("nlayers"), "nlayers"), as_int(e.get("nx"), "nx"),
               as_int(e.get("ny"), "ny"), as_int(e.get("dtype"), "dtype")};
    const Value* p = e.get("p");
    if (!p || p->kind != Value::ARR || p->arr.size() != 3) throw std::runtime_error("schedule table: 'p' must hold 3 integers");
    Sched s;
    for (int j = 0; j < 3; ++j) s.p[j] = as_int(&p->arr[j], "p");
    const Value* sp = e.get("split");
    if (sp) {
      if (sp->kind != Value::ARR || (sp->arr.size() != 2 && sp->arr.size() != 3))
        throw std::runtime_error("schedule table: 'split' must hold 2 or 3 integers");
      for (int j = 0; j < (int)sp->arr.size(); ++j) {
        s.split[j] = as_int(&sp->arr[j], "split");
        if (s.split[j] < 0 || s.split[j] > (j == 1 ? 2 : 1))
          throw std::runtime_error("schedule table: 'split' entries are 0 or 1 (0, 1 or 2 for K2)");
      }
      if ((s.split[0] || s.split[1] || s.split[2]) && k[1] != 1)
        throw std::runtime_error("schedule table: workgroup splits are for single-layer problems");
    }
    const Value* us = e.get("step_us");
    if (us && us->kind == Value::NUM) s.step_us = us->num;
    // K1 and K3 transform columns (length ny), K2 rows (length nx)
    if (!offered(k[3], s.p[0]) || !offered(k[2], s.p[1]) || !offered(k[3], s.p[2]))
      throw std::runtime_error("schedule table: entry for " + std::to_string(k[2]) + "x" + std::to_string(k[3]) +
                               " names a schedule this build does not offer");
    got.emplace_back(k, s);
  }
  for (auto& kv : got) sched_store(kv.first, kv.second);
  return (int)got.size();
}

}  // namespace gf
