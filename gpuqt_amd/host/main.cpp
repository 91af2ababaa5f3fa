// main.cpp -- the lsqt_gpu program: `lsqt_gpu <driver-file>`, one input
// directory per non-empty line (src/main.cu:34-56).
#include "lsqt_host.h"

int main(int argc, char* argv[])
{
  return lsqt_host_main(argc, argv);
}
