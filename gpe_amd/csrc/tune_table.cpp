// tune_table.cpp -- reader of the persisted launch-configuration table.
//
// The table is a JSON object keyed by problem ("NXxNYxNZ/oORDER/PREC/SCHEME/
// bBCS[/fused]"); each value maps a stage-kind name to [variant, zchunk]:
//
//   {"256x256x256/o4/fp64/rk4/b2": {"rk4_1": [3, 0], "rk4_2": [1, 64], ...}, ...}
//
// The autotuner's results are written in this form by the Python host code;
// this reader lets C and Fortran callers apply them.  It accepts any JSON
// document (strings with escapes, nested values, numbers) and extracts only
// the object stored under one key.
#include <cctype>
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "host_common.h"

namespace gpe {
namespace {

struct Cursor {
  const char *s, *end;
  std::string err;
  void ws() {
    while (s < end && std::isspace(static_cast<unsigned char>(*s))) ++s;
  }
  bool expect(char ch) {
    ws();
    if (s < end && *s == ch) {
      ++s;
      return true;
    }
    if (err.empty()) err = std::string("expected '") + ch + "'";
    return false;
  }
  bool peek(char ch) {
    ws();
    return s < end && *s == ch;
  }
  bool string(std::string &out) {
    if (!expect('"')) return false;
    out.clear();
    while (s < end && *s != '"') {
      if (*s == '\\') {
        if (++s >= end) break;
      }
      out.push_back(*s++);
    }
    if (s >= end) {
      err = "unterminated string";
      return false;
    }
    ++s;
    return true;
  }
  bool integer(long &v) {
    ws();
    char *e = nullptr;
    errno = 0;
    v = std::strtol(s, &e, 10);
    if (e == s || errno != 0 || e > end) {
      err = "expected an integer";
      return false;
    }
    s = e;
    return true;
  }
  // Skip one value of any type (nested at most kMaxDepth deep, so a
  // malformed file cannot exhaust the stack).
  static constexpr int kMaxDepth = 64;
  bool skip(int depth = 0) {
    ws();
    if (s >= end) {
      err = "unexpected end";
      return false;
    }
    if (depth > kMaxDepth) {
      err = "values nested too deep";
      return false;
    }
    if (*s == '"') {
      std::string tmp;
      return string(tmp);
    }
    if (*s == '{' || *s == '[') {
      const char close = *s == '{' ? '}' : ']';
      const bool obj = *s == '{';
      ++s;
      if (peek(close)) {
        ++s;
        return true;
      }
      for (;;) {
        if (obj) {
          std::string k;
          if (!string(k) || !expect(':')) return false;
        }
        if (!skip(depth + 1)) return false;
        if (peek(',')) {
          ++s;
          continue;
        }
        return expect(close);
      }
    }
    // number, true, false, null
    const char *b = s;
    while (s < end && (std::isalnum(static_cast<unsigned char>(*s)) || *s == '-' || *s == '+' || *s == '.')) ++s;
    if (s == b) {
      err = "unexpected character";
      return false;
    }
    return true;
  }
  // {"name": [v, z], ...}
  bool entry(std::vector<TuneItem> &out) {
    if (!expect('{')) return false;
    if (peek('}')) {
      ++s;
      return true;
    }
    for (;;) {
      TuneItem it;
      long v = 0, z = 0;
      if (!string(it.name) || !expect(':') || !expect('[') || !integer(v) || !expect(',') || !integer(z) ||
          !expect(']'))
        return false;
      // values that do not fit an int32 are invalid, not wrapped
      if (v < INT32_MIN || v > INT32_MAX || z < INT32_MIN || z > INT32_MAX) {
        err = "value out of range in " + it.name;
        return false;
      }
      it.variant = static_cast<int>(v);
      it.zchunk = static_cast<int>(z);
      out.push_back(it);
      if (peek(',')) {
        ++s;
        continue;
      }
      return expect('}');
    }
  }
};

}  // namespace

int read_tune_table(const char *path, const char *key, std::vector<TuneItem> &out, bool &found) {
  out.clear();
  found = false;
  FILE *fh = std::fopen(path, "rb");
  if (!fh) {
    if (errno == ENOENT) return GPE_OK;  // no table: nothing stored
    return fail(GPE_EIO, "%s: %s", path, std::strerror(errno));
  }
  std::string text;
  char buf[65536];
  size_t n;
  while ((n = std::fread(buf, 1, sizeof buf, fh)) > 0) text.append(buf, n);
  const bool bad = std::ferror(fh) != 0;
  std::fclose(fh);
  if (bad) return fail(GPE_EIO, "%s: read error", path);

  Cursor c{text.data(), text.data() + text.size(), {}};
  if (!c.expect('{')) return fail(GPE_EIO, "%s: %s", path, c.err.c_str());
  if (c.peek('}')) return GPE_OK;
  for (;;) {
    std::string k;
    if (!c.string(k) || !c.expect(':')) return fail(GPE_EIO, "%s: %s", path, c.err.c_str());
    if (k == key) {
      if (!c.entry(out)) return fail(GPE_EIO, "%s: entry %s: %s", path, key, c.err.c_str());
      found = true;
      return GPE_OK;
    }
    if (!c.skip()) return fail(GPE_EIO, "%s: %s", path, c.err.c_str());
    if (c.peek(',')) {
      ++c.s;
      continue;
    }
    if (!c.expect('}')) return fail(GPE_EIO, "%s: %s", path, c.err.c_str());
    return GPE_OK;
  }
}

}  // namespace gpe
