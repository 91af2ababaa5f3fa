// text_reader.h -- a fast whitespace-separated number reader for the large
// input files (neighbor.in, hopping.in, position.in, potential.in of the
// general model, src/model_general.cu:32-185, which reads them with
// std::ifstream >>).
//
// The whole file is mapped into memory and numbers are converted with
// std::from_chars, which rounds correctly like the stream's strtod, so a
// well-formed file gives the same values bit for bit.  The extraction
// operators keep the stream's failure rules: the first failed extraction
// stores 0 and sets the failure state, later extractions leave their targets
// unchanged.  A token outside the plain grammar ([+-]digits[.digits]
// [(e|E)[+-]digits], or a word for strings) marks the reader `unusual`:
// callers then parse the file again with std::ifstream, so that every input
// the stream accepts or rejects is handled exactly as the stream does.
#pragma once

#include <charconv>
#include <cstddef>
#include <cstring>
#include <string>
#include <system_error>
#include <type_traits>

namespace lsqt_host {

class TextReader {
 public:
  explicit TextReader(const std::string& path);
  ~TextReader();
  TextReader(const TextReader&) = delete;
  TextReader& operator=(const TextReader&) = delete;

  bool is_open() const { return opened_; }
  bool fail() const { return fail_; }
  bool unusual() const { return unusual_; }

  TextReader& operator>>(int& v) { return number(v); }
  TextReader& operator>>(double& v) { return number(v); }
  TextReader& operator>>(std::string& s)
  {
    if (fail_)
      return *this;
    if (!skip_space()) {
      s.clear();
      fail_ = true;
      return *this;
    }
    const char* b = p_;
    while (p_ < end_ && !space(*p_))
      ++p_;
    s.assign(b, p_);
    return *this;
  }

 private:
  static bool space(char c)
  {
    return c == ' ' || c == '\t' || c == '\n' || c == '\r' || c == '\v' || c == '\f';
  }
  bool skip_space()
  {
    while (p_ < end_ && space(*p_))
      ++p_;
    return p_ < end_;
  }
  // length of a plain-grammar number token at p_, 0 if none
  size_t plain_token(bool integer) const
  {
    const char* q = p_;
    if (q < end_ && (*q == '+' || *q == '-'))
      ++q;
    const char* d0 = q;
    while (q < end_ && *q >= '0' && *q <= '9')
      ++q;
    bool digits = q > d0;
    if (!integer) {
      if (q < end_ && *q == '.') {
        ++q;
        const char* f0 = q;
        while (q < end_ && *q >= '0' && *q <= '9')
          ++q;
        digits = digits || q > f0;
      }
      if (digits && q < end_ && (*q == 'e' || *q == 'E')) {
        const char* e = q + 1;
        if (e < end_ && (*e == '+' || *e == '-'))
          ++e;
        const char* e0 = e;
        while (e < end_ && *e >= '0' && *e <= '9')
          ++e;
        if (e == e0)
          return 0;  // "1e" and the like: leave to the stream
        q = e;
      }
    }
    if (!digits || (q < end_ && !space(*q)))
      return 0;
    return (size_t)(q - p_);
  }
  template <class T>
  TextReader& number(T& v)
  {
    if (fail_)
      return *this;
    if (!skip_space()) {  // end of file: the stream stores 0 and fails
      v = 0;
      fail_ = true;
      return *this;
    }
    const size_t len = plain_token(std::is_integral<T>::value);
    if (len == 0) {
      unusual_ = true;
      v = 0;
      fail_ = true;
      return *this;
    }
    const char* b = p_;
    if (*b == '+')
      ++b;
    T x{};
    auto r = std::from_chars(b, p_ + len, x);
    if (r.ec != std::errc() || r.ptr != p_ + len) {
      unusual_ = true;  // out of range and the like: the stream decides
      v = 0;
      fail_ = true;
      return *this;
    }
    v = x;
    p_ += len;
    return *this;
  }

  bool opened_ = false;
  bool fail_ = false;
  bool unusual_ = false;
  char* data_ = nullptr;
  size_t size_ = 0;
  const char* p_ = nullptr;
  const char* end_ = nullptr;
};

// Writes numbers in the shortest form that reads back to the same bits.
inline void append_number(std::string& out, double v)
{
  char buf[64];
  auto r = std::to_chars(buf, buf + sizeof(buf), v);
  out.append(buf, r.ptr);
}

inline void append_number(std::string& out, int v)
{
  char buf[16];
  auto r = std::to_chars(buf, buf + sizeof(buf), v);
  out.append(buf, r.ptr);
}

}  // namespace lsqt_host
