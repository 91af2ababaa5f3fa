// kernels_build.hip -- Hamiltonian construction on the device: the packed
// slot words of a lattice model (src/model_lattice.cu:148-163, 342-386).
//
// One thread per pristine site p = ((nx*Ny + ny)*Nz + nz)*N_orbital + m (x
// slowest), its hoppings taken in file order, those leaving an open boundary
// skipped (so later ones move up a slot), the neighbour cell wrapped once
// across periodic boundaries.  With vacancies (src/model_lattice.cu:70-146)
// the site and its neighbours take their indices after the vacancies were
// removed, vacancies write nothing, and neighbours that are vacancies are
// dropped with the later slots moving up.  The class of a slot is that of its
// (orbital, hopping) pair, whose value and transport offset the host
// tabulates; unused slots hold the empty marker.
#include "lsqt_device.h"
#include "lsqt_launch.h"

namespace lsqt {

__global__ __launch_bounds__(kBlock) void k_lattice_words(LatticeDev L, uint32_t* __restrict__ words,
                                                          int stride, int idx_bits, int Kmax,
                                                          int n_pristine)
{
  const int p = blockIdx.x * kBlock + threadIdx.x;
  if (p >= n_pristine)
    return;
  const int own = L.site_index ? L.site_index[p] : p;
  if (own < 0)
    return;  // a vacancy
  const int site = L.offset + own;
  const int m = p % L.orbitals;
  const int cell = p / L.orbitals;
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
    const int j1 = ((c2[0] * L.cells[1] + c2[1]) * L.cells[2] + c2[2]) * L.orbitals + h.w;
    const int j = L.site_index ? L.site_index[j1] : j1;
    if (j < 0)
      continue;  // the neighbour is a vacancy: later slots move up
    words[(size_t)count * stride + site] = ((uint32_t)L.cls[k] << idx_bits) |
                                           (uint32_t)(L.offset + j);
    ++count;
  }
  for (; count < Kmax; ++count)
    words[(size_t)count * stride + site] = kEmptySlot;
}

void launch_lattice_words(const LatticeDev& L, uint32_t* words, int stride, int idx_bits,
                          int Kmax, int n_pristine, hipStream_t st)
{
  k_lattice_words<<<(n_pristine + kBlock - 1) / kBlock, kBlock, 0, st>>>(L, words, stride,
                                                                         idx_bits, Kmax,
                                                                         n_pristine);
}

}  // namespace lsqt
