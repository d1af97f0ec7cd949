// Text formatting shared by the output writers (host code only, no HIP).
#pragma once

#include <charconv>
#include <cstdio>
#include <cstdlib>
#include <cstring>

namespace mw {

// Python's repr of a float (shortest round-trip digits; positional for decimal
// exponents -4..15, else d.ddde[+-]XX): what numpy's tofile(sep=" ") and "%s"
// print for a float64.  Writes at most 32 characters to out, returns the count.
inline int py_repr_chars(double v, char* out) {
  if (v != v) return (int)(stpcpy(out, "nan") - out);
  if (v > 1.7976931348623157e308) return (int)(stpcpy(out, "inf") - out);
  if (v < -1.7976931348623157e308) return (int)(stpcpy(out, "-inf") - out);
  char buf[40];
  char* e = std::to_chars(buf, buf + sizeof(buf), v, std::chars_format::scientific).ptr;
  *e = 0;
  char* o = out;
  const char* p = buf;
  if (*p == '-') *o++ = *p++;
  char digits[24];
  int nd = 0;
  for (; *p != 'e'; p++)
    if (*p != '.') digits[nd++] = *p;
  const int ex = atoi(p + 1);
  if (ex >= -4 && ex < 16) {
    if (ex >= 0) {
      for (int k = 0; k <= ex; k++) *o++ = k < nd ? digits[k] : '0';
      *o++ = '.';
      if (nd > ex + 1)
        for (int k = ex + 1; k < nd; k++) *o++ = digits[k];
      else
        *o++ = '0';
    } else {
      *o++ = '0';
      *o++ = '.';
      for (int k = 0; k < -ex - 1; k++) *o++ = '0';
      for (int k = 0; k < nd; k++) *o++ = digits[k];
    }
  } else {
    *o++ = digits[0];
    if (nd > 1) {
      *o++ = '.';
      for (int k = 1; k < nd; k++) *o++ = digits[k];
    }
    o += snprintf(o, 8, "e%c%02d", ex < 0 ? '-' : '+', ex < 0 ? -ex : ex);
  }
  *o = 0;
  return (int)(o - out);
}

// printf's "%.<prec>e" (correctly rounded, at least two exponent digits),
// through std::to_chars.  Writes at most 32 characters, returns the count.
inline int sci_chars(double v, int prec, char* out) {
  char* e = std::to_chars(out, out + 32, v, std::chars_format::scientific, prec).ptr;
  *e = 0;
  return (int)(e - out);
}

// printf's "%ld"
inline int int_chars(long v, char* out) {
  char* e = std::to_chars(out, out + 24, v).ptr;
  *e = 0;
  return (int)(e - out);
}

}  // namespace mw
