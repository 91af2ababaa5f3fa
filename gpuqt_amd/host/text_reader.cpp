// text_reader.cpp -- file mapping of TextReader.
#include "text_reader.h"

#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

namespace lsqt_host {

TextReader::TextReader(const std::string& path)
{
  int fd = ::open(path.c_str(), O_RDONLY);
  if (fd < 0)
    return;
  struct stat st;
  if (::fstat(fd, &st) != 0 || !S_ISREG(st.st_mode)) {
    ::close(fd);
    return;
  }
  opened_ = true;
  size_ = (size_t)st.st_size;
  if (size_ > 0) {
    void* m = ::mmap(nullptr, size_, PROT_READ, MAP_PRIVATE, fd, 0);
    if (m == MAP_FAILED) {
      opened_ = false;
      size_ = 0;
    } else {
      data_ = (char*)m;
      ::madvise(m, size_, MADV_SEQUENTIAL);
    }
  }
  ::close(fd);
  p_ = data_;
  end_ = data_ ? data_ + size_ : nullptr;
}

TextReader::~TextReader()
{
  if (data_)
    ::munmap(data_, size_);
}

}  // namespace lsqt_host
