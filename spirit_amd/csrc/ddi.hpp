// Device-side description of one DDI FFT convolution plan and its launchers.
#pragma once

#include "device_types.hpp"

namespace spirit_hip
{

struct DdiDev
{
    int Na, Nb, Nc, nb; // lattice cells and basis atoms
    int Pa, Pb, Pc, Ha; // padded grid, Ha = Pa/2 + 1
    int n_inter;        // distinct sublattice pairs (Hamiltonian_Heisenberg.cpp:1545-1554)
    int pad;
    int inter[MAX_BASIS * MAX_BASIS]; // inter_sublattice_lookup[b1 + b2*nb]
    double mu[MAX_BASIS];
    double M;            // Pa*Pb*Pc (the unnormalised FFT scale)
    FftLen La, Lb, Lc;
    double2 * X;         // spin spectrum [3nb][Nc][Pb][Ha]
    const double2 * Dh;  // kernel spectrum [Pb][Ha][Pc][6*n_inter]
};

struct DipoleArgs
{
    int Na, Nb, Nc, nb;
    int Pa, Pb, Pc, pad;
    int img[3];
    int pad2;
    double lattice_constant;
    double bravais[3][3];
    double cell_atoms[MAX_BASIS][3];
};

int fft_lines_per_block( const FftLen & L );

// h = mu_i * (D * (mu s))_i / M for the spins in `spins` (SoA, nos), written to h (SoA).
// evc0/evc1 (nullable) are recorded around the convolution pass (pass C) for profiling.
void ddi_launch_field( const DdiDev & d, const double * spins, double * h, int nos, hipStream_t st,
                       hipEvent_t evc0 = nullptr, hipEvent_t evc1 = nullptr );
// Setup of the kernel spectrum.
void ddi_launch_kernel_spectrum( const DdiDev & d, const DipoleArgs & g, double * work, double2 * work2,
                                 double2 * Dh, hipStream_t st );

} // namespace spirit_hip
