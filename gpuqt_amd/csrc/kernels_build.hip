// kernels_build.hip -- Hamiltonian construction on the device: the packed
// slot words of a lattice model (src/model_lattice.cu:148-163, 342-386).
//
// One thread per site: site n = ((nx*Ny + ny)*Nz + nz)*N_orbital + m (x
// slowest), its hoppings taken in file order, those leaving an open boundary
// skipped (so later ones move up a slot), the neighbour cell wrapped once
// across periodic boundaries.  The class of a slot is the (orbital, hopping)
// pair m*K + i, whose value (and transport offset) the host tabulates; unused
// slots hold the empty marker.
#include "lsqt_device.h"
#include "lsqt_launch.h"

namespace lsqt {

__global__ __launch_bounds__(kBlock) void k_lattice_words(LatticeDev L, uint32_t* __restrict__ words,
                                                          int stride, int idx_bits, int n)
{
  const int site = blockIdx.x * kBlock + threadIdx.x;
  if (site >= n)
    return;
  const int m = site % L.orbitals;
  const int cell = site / L.orbitals;
  const int nz1 = cell % L.cells[2];
  const int ny1 = (cell / L.cells[2]) % L.cells[1];
  const int nx1 = cell / (L.cells[2] * L.cells[1]);
  int count = 0;
  for (int i = 0; i < L.hop_count[m]; ++i) {
    const int k = m * L.K + i;
    const int4 h = L.hop[k];  // cell offsets and target orbital
    int c2[3] = {nx1 + h.x, ny1 + h.y, nz1 + h.z};
    bool skip = false;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      if (c2[d] < 0 || c2[d] >= L.cells[d]) {
        if (!L.pbc[d])
          skip = true;
        else
          c2[d] += c2[d] < 0 ? L.cells[d] : -L.cells[d];
      }
    }
    if (skip)
      continue;
    const int j = ((c2[0] * L.cells[1] + c2[1]) * L.cells[2] + c2[2]) * L.orbitals + h.w;
    words[(size_t)count * stride + site] = ((uint32_t)k << idx_bits) | (uint32_t)j;
    ++count;
  }
  for (; count < L.K; ++count)
    words[(size_t)count * stride + site] = kEmptySlot;
}

void launch_lattice_words(const LatticeDev& L, uint32_t* words, int stride, int idx_bits, int n,
                          hipStream_t st)
{
  k_lattice_words<<<(n + kBlock - 1) / kBlock, kBlock, 0, st>>>(L, words, stride, idx_bits, n);
}

}  // namespace lsqt
