// general_io.cpp -- writer of the general-model input files.
#include "general_io.h"

#include <cstdio>
#include <string>

#include "common.h"
#include "text_reader.h"

namespace lsqt_host {
namespace {

class FileOut {
 public:
  explicit FileOut(const std::string& path) : path_(path), f_(std::fopen(path.c_str(), "wb"))
  {
    if (!f_)
      throw InputError("Error: cannot open " + path + " for writing");
    buf_.reserve(kFlush + 4096);
  }
  ~FileOut()
  {
    if (f_)
      std::fclose(f_);
  }
  std::string& buf() { return buf_; }
  void maybe_flush()
  {
    if (buf_.size() >= kFlush)
      flush();
  }
  void close()
  {
    flush();
    if (std::fclose(f_) != 0) {
      f_ = nullptr;
      throw InputError("Error: cannot write " + path_);
    }
    f_ = nullptr;
  }

 private:
  void flush()
  {
    if (!buf_.empty() && std::fwrite(buf_.data(), 1, buf_.size(), f_) != buf_.size())
      throw InputError("Error: cannot write " + path_);
    buf_.clear();
  }
  static constexpr size_t kFlush = 1 << 22;
  std::string path_;
  std::FILE* f_;
  std::string buf_;
};

}  // namespace

void write_general(const std::string& dir, int n, int K, const int* nn, const int* nl,
                   const double* x, double box, double volume, const double* pot,
                   const double* hr, const double* hi)
{
  {
    FileOut out(dir + "/neighbor.in");
    std::string& b = out.buf();
    append_number(b, n);
    b += ' ';
    append_number(b, K);
    b += '\n';
    for (int i = 0; i < n; ++i) {
      append_number(b, nn[i]);
      for (int m = 0; m < nn[i]; ++m) {
        b += ' ';
        append_number(b, nl[(size_t)i * K + m]);
      }
      b += '\n';
      out.maybe_flush();
    }
    out.close();
  }
  {
    FileOut out(dir + "/position.in");
    std::string& b = out.buf();
    append_number(b, box);
    b += ' ';
    append_number(b, volume);
    b += '\n';
    for (int i = 0; i < n; ++i) {
      append_number(b, x[i]);
      b += '\n';
      out.maybe_flush();
    }
    out.close();
  }
  if (pot) {
    FileOut out(dir + "/potential.in");
    std::string& b = out.buf();
    for (int i = 0; i < n; ++i) {
      append_number(b, pot[i]);
      b += '\n';
      out.maybe_flush();
    }
    out.close();
  }
  {
    FileOut out(dir + "/hopping.in");
    std::string& b = out.buf();
    b += hi ? "complex\n" : "real\n";
    for (int i = 0; i < n; ++i) {
      for (int m = 0; m < nn[i]; ++m) {
        const size_t q = (size_t)i * K + m;
        if (m)
          b += ' ';
        append_number(b, hr[q]);
        if (hi) {
          b += ' ';
          append_number(b, hi[q]);
        }
      }
      b += '\n';
      out.maybe_flush();
    }
    out.close();
  }
}

}  // namespace lsqt_host
