// Device description of the real-space dipole-dipole methods (cutoff pairs, direct sum).
#pragma once

#include "device_types.hpp"

namespace spirit_hip
{

enum DdiRsMethod
{
    DDI_RS_CUTOFF = 1, // Gradient_DDI_Cutoff, Hamiltonian_Heisenberg.cpp:880-918
    DDI_RS_DIRECT = 2, // Gradient_DDI_Direct, Hamiltonian_Heisenberg.cpp:1016-1071
};

// One cutoff pair (Neighbours.cpp:158-246): basis atoms i -> j, translation, the prefactor
// c = mu_0 mu_B^2 / (4 pi 1e-30) / |r_ij|^3 and the unit vector n = r_ij / |r_ij|
// (DDI_from_Pair, Neighbours.cpp:288-307).
struct DdiPairDev
{
    int i, j, da, db, dc, pad;
    double c;
    double n[3];
};

struct DdiRsDev
{
    int method;
    int nos, nb, Na, Nb, Nc;
    int bc[3];
    int img[3]; // periodic images of the direct sum (0 along open axes)
    int n_pairs;
    double mult; // mu_0 mu_B^2 / (4 pi 1e-30)
    double lc;   // lattice constant
    double bv[3][3];
    double mu[MAX_BASIS];
    const double * pos;        // positions, SoA [x...][y...][z...] (direct sum)
    const DdiPairDev * pairs;  // n_pairs (cutoff)
    FastDiv div_nb, div_Na, div_Nb;
};

// h (SoA) = the dipolar field of `spins` (SoA): g_i = -h_i
void ddi_rs_launch_field( const DdiRsDev & D, const double * spins, double * h, hipStream_t st );

} // namespace spirit_hip
