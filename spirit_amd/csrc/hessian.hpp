// Embedding Hessian of the Heisenberg Hamiltonian on the GPU (Hamiltonian_Heisenberg::Hessian).
#pragma once

#include "host_core.hpp"

namespace spirit_hip
{

// Dense 3N x 3N embedding Hessian (row-major, index 3*ispin + component) into host memory,
// Hamiltonian_Heisenberg.cpp:1119-1245.
void dense_hessian( const Geometry & g, const Hamiltonian & h, double * out );

// out = H v for the same matrix (AoS host arrays of 3N), without forming it.
void hessian_vector_product( const Geometry & g, const Hamiltonian & h, const double * v, double * out );

} // namespace spirit_hip
