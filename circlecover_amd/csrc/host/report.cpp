// Interval text: the reference's writer (algcuda/interval.hpp:294-303) and
// reader (algcuda/interval.hpp:305-351), so that reports can be read back
// and compared bit for bit.  Host-only; no device context.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <sstream>
#include <string>

#include "circlecover_amd.h"

namespace {

bool is_space(char c) { return c == ' ' || c == '\t' || c == '\n' || c == '\r' || c == '\f' || c == '\v'; }
bool is_digit(char c) { return c >= '0' && c <= '9'; }

// One floating-point number as an iostream extraction of a double reads it
// in the "C" locale: leading white space skipped, then the longest prefix of
// [+-] digits [. digits] [(e|E) [+-] digits] with at least one mantissa
// digit, converted with correct rounding.  No "inf", "nan" or hex forms
// ("0x10" reads as 0, stopping at the x).
bool read_number(const char* s, size_t n, size_t& pos, double& out) {
  while (pos < n && is_space(s[pos])) ++pos;
  size_t p = pos;
  std::string tok;
  if (p < n && (s[p] == '+' || s[p] == '-')) tok += s[p++];
  size_t mant = 0;
  while (p < n && is_digit(s[p])) tok += s[p++], ++mant;
  if (p < n && s[p] == '.') {
    tok += s[p++];
    while (p < n && is_digit(s[p])) tok += s[p++], ++mant;
  }
  if (mant == 0) return false;
  if (p < n && (s[p] == 'e' || s[p] == 'E')) {
    size_t q = p + 1;
    std::string ex = "e";
    if (q < n && (s[q] == '+' || s[q] == '-')) ex += s[q++];
    size_t ed = 0;
    while (q < n && is_digit(s[q])) ex += s[q++], ++ed;
    if (ed == 0) return false;  // an exponent mark without digits fails the extraction
    tok += ex;
    p = q;
  }
  char* end = nullptr;
  const double v = std::strtod(tok.c_str(), &end);
  // an overflow fails the extraction; an underflow reads as the rounded
  // (subnormal or zero) value
  if (end != tok.c_str() + tok.size() || v == HUGE_VAL || v == -HUGE_VAL) return false;
  out = v;
  pos = p;
  return true;
}

// One non-space character, as `input >> c` reads it.
bool read_char(const char* s, size_t n, size_t& pos, char& c) {
  while (pos < n && is_space(s[pos])) ++pos;
  if (pos >= n) return false;
  c = s[pos++];
  return true;
}

}  // namespace

extern "C" {

int cc_format_interval(const cc_iv* x, char* buf, size_t cap, size_t* len) {
  if (!x || !len) return CC_E_ARG;
  std::ostringstream o;
  o.flags(std::ios::right);
  o.precision(19);
  o << '[';
  o.width(26);
  o << x->lb << ", ";
  o.width(26);
  o << x->ub << ']';
  const std::string t = o.str();
  *len = t.size();
  if (!buf || cap <= t.size()) return CC_E_OVERFLOW;
  std::memcpy(buf, t.c_str(), t.size() + 1);
  return CC_OK;
}

int cc_parse_interval(const char* text, size_t n, cc_iv* out, size_t* consumed) {
  if (!text || !out) return CC_E_ARG;
  size_t pos = 0;
  char c;
  if (!read_char(text, n, pos, c)) return CC_E_ARG;
  cc_iv v;
  if (c == '[') {
    char sep, end;
    if (!read_number(text, n, pos, v.lb) || !read_char(text, n, pos, sep)) return CC_E_ARG;
    if (sep != ',' && sep != ';') return CC_E_ARG;
    if (!read_number(text, n, pos, v.ub) || !read_char(text, n, pos, end)) return CC_E_ARG;
    if (end != ']') return CC_E_ARG;
  } else {
    --pos;  // the character goes back and a scalar x reads as [x, x]
    if (!read_number(text, n, pos, v.lb)) return CC_E_ARG;
    v.ub = v.lb;
  }
  *out = v;
  if (consumed) *consumed = pos;
  return CC_OK;
}

}  // extern "C"
