// hamiltonian.cpp -- Hamiltonian over the C ABI.
#include "hamiltonian.h"

#include "lsqt_hip.h"
#include "model.h"
#include "vector.h"

namespace lsqt_host {

lsqt_lattice abi_lattice(const LatticeSpec& L)
{
  lsqt_lattice a{};
  for (int d = 0; d < 3; ++d) {
    a.cells[d] = L.N_cell[d];
    a.pbc[d] = L.pbc[d];
    a.lattice_constant[d] = L.lattice_constant[d];
  }
  a.transport = L.transport_direction;
  a.orbitals = L.N_orbital;
  a.orbital_x = L.x_cell.data();
  a.max_neighbor = L.max_neighbor;
  a.hop_count = L.number_of_hoppings.data();
  a.hop_cell = L.hop_cell.data();
  a.hop_orbital = L.hop_orbital.data();
  a.hop_re = L.hop_re.data();
  a.hop_im = L.hop_im.data();
  return a;
}

}  // namespace lsqt_host

Hamiltonian::Hamiltonian(Model& model)
{
  if (const LatticeSpec* L = model.lattice()) {
    // a lattice model whose neighbour lists were not expanded on the host:
    // expanded on the GPU (the same arrays, so the same results bit for bit)
    const lsqt_lattice lat = lsqt_host::abi_lattice(*L);
    const int* index = L->site_index.empty() ? nullptr : L->site_index.data();
    LSQT_CHECK(lsqt_ham_create_lattices(lsqt_host::default_context(), 1, &lat,
                                        index ? &index : nullptr, model.potential,
                                        model.energy_max, &h));
  } else {
    LSQT_CHECK(lsqt_ham_create(lsqt_host::default_context(), model.number_of_atoms,
                               model.max_neighbor, model.neighbor_number, model.neighbor_list,
                               model.potential, model.hopping_real, model.hopping_imag,
                               model.xx, model.energy_max, LSQT_LAYOUT_AUTO, &h));
  }
  model.release_hamiltonian_arrays();
}

Hamiltonian::~Hamiltonian()
{
  lsqt_ham_destroy(h);
}

void Hamiltonian::apply(Vector& input, Vector& output)
{
  LSQT_CHECK(lsqt_apply(h, input.handle(), output.handle()));
}

void Hamiltonian::apply_commutator(Vector& input, Vector& output)
{
  LSQT_CHECK(lsqt_apply_commutator(h, input.handle(), output.handle()));
}

void Hamiltonian::apply_current(Vector& input, Vector& output)
{
  LSQT_CHECK(lsqt_apply_current(h, input.handle(), output.handle()));
}

void Hamiltonian::kernel_polynomial(Vector& s0, Vector& s1, Vector& s2)
{
  LSQT_CHECK(lsqt_kernel_polynomial(h, s0.handle(), s1.handle(), s2.handle()));
}

void Hamiltonian::chebyshev_01(Vector& s0, Vector& s1, Vector& s, real b0, real b1, int dir)
{
  LSQT_CHECK(lsqt_chebyshev_01(h, s0.handle(), s1.handle(), s.handle(), b0, b1, dir));
}

void Hamiltonian::chebyshev_2(Vector& s0, Vector& s1, Vector& s2, Vector& s, real bm, int label)
{
  LSQT_CHECK(lsqt_chebyshev_2(h, s0.handle(), s1.handle(), s2.handle(), s.handle(), bm, label));
}

void Hamiltonian::chebyshev_1x(Vector& s1x, Vector& s, real b1)
{
  LSQT_CHECK(lsqt_chebyshev_1x(h, s1x.handle(), s.handle(), b1));
}

void Hamiltonian::chebyshev_2x(Vector& s0, Vector& s0x, Vector& s1, Vector& s1x, Vector& s2,
                               Vector& s2x, Vector& s, real bm, int label)
{
  LSQT_CHECK(lsqt_chebyshev_2x(h, s0.handle(), s0x.handle(), s1.handle(), s1x.handle(),
                               s2.handle(), s2x.handle(), s.handle(), bm, label));
}
