// Text loader for relation files: the reference's pipe-delimited format (main/main.cpp:49-109).
// One tuple per line, fields separated by '|', blank lines and lines starting with '#' skipped,
// each field parsed like std::stoi (leading blanks, optional sign, decimal digits; FromString,
// types/FromString.hpp:23-30).  The reference grows its buffers only after writing past them
// when a file has exactly `capacity` lines (main/main.cpp:73-80); here the caller's capacity is
// checked before every row.
#include "quickfoil_amd.h"

#include <cerrno>
#include <climits>
#include <cstdint>
#include <cstdlib>

namespace {

// Parses one field [p, end) like std::stoi: returns false on no digits or int32 overflow.
bool parse_field(const char* p, const char* end, int32_t* out) {
  while (p < end && (*p == ' ' || *p == '\t' || *p == '\r' || *p == '\v' || *p == '\f')) ++p;
  bool neg = false;
  if (p < end && (*p == '+' || *p == '-')) {
    neg = *p == '-';
    ++p;
  }
  if (p >= end || *p < '0' || *p > '9') return false;
  int64_t v = 0;
  while (p < end && *p >= '0' && *p <= '9') {
    v = v * 10 + (*p - '0');
    if (v > (int64_t)INT32_MAX + 1) return false;
    ++p;
  }
  if (neg) v = -v;
  if (v < INT32_MIN || v > INT32_MAX) return false;
  *out = (int32_t)v;
  return true;
}

}  // namespace

extern "C" int qf_parse_relation_text(const char* text, int64_t length, int num_fields, const int* keep,
                                      int32_t* const* h_columns, int64_t capacity, int64_t* h_num_rows,
                                      int64_t* h_error_line) {
  if (!text || length < 0 || num_fields <= 0 || !h_num_rows) return QF_ERR_INVALID;
  int64_t rows = 0, line_no = 0;
  const char* p = text;
  const char* end = text + length;
  if (h_error_line) *h_error_line = -1;
  while (p < end) {
    const char* eol = p;
    while (eol < end && *eol != '\n') ++eol;
    ++line_no;
    if (eol > p && *p != '#' && !(eol - p == 1 && *p == '\r')) {
      int field = 0, col = 0;
      const char* f = p;
      bool ok = true;
      while (true) {
        const char* fe = f;
        while (fe < eol && *fe != '|') ++fe;
        if (field >= num_fields) {
          ok = false;
          break;
        }
        if (!keep || keep[field]) {
          int32_t v;
          if (!parse_field(f, fe, &v)) {
            ok = false;
            break;
          }
          if (h_columns) {
            if (rows >= capacity) {
              *h_num_rows = rows;
              return QF_ERR_OVERFLOW;
            }
            h_columns[col][rows] = v;
          }
          ++col;
        }
        ++field;
        if (fe >= eol) break;
        f = fe + 1;
      }
      if (!ok || field != num_fields) {
        if (h_error_line) *h_error_line = line_no;
        *h_num_rows = rows;
        return QF_ERR_INVALID;
      }
      ++rows;
    }
    p = eol < end ? eol + 1 : end;
  }
  *h_num_rows = rows;
  return QF_OK;
}
