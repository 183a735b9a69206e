// Device-side description of one DDI FFT convolution plan and its launchers.
#pragma once

#include "device_types.hpp"

#include <utility>
#include <vector>

namespace spirit_hip
{

struct DdiDev
{
    int Na, Nb, Nc, nb; // lattice cells and basis atoms
    int Pa, Pb, Pc, Ha; // padded grid, Ha = Pa/2 + 1
    int Hs;             // row stride of X: Ha rounded up to a multiple of 8 (128-byte rows)
    int n_inter;        // distinct sublattice pairs (Hamiltonian_Heisenberg.cpp:1545-1554)
    int inter[MAX_BASIS * MAX_BASIS]; // inter_sublattice_lookup[b1 + b2*nb]
    double mu[MAX_BASIS];
    double M; // Pa*Pb*Pc (the unnormalised FFT scale)
    FftLen La, Lb, Lc;
    double2 * X; // spin spectrum [3nb][Nc][Pb][Hs]
    // Kernel spectrum, stored [kb'][ka][kc'][6*n_inter] with kb' < KB, kc' < KC.
    // dh_real: entries are real (D even in r); sym_b/sym_c: only kb <= Pb/2 / kc <= Pc/2 are
    // stored and the mirror is component-wise +-1 (sig_b/sig_c) times the stored value.
    const double2 * Dh;
    const double * Dh_r;
    int dh_real, sym_b, sym_c;
    int KB, KC;
    int sig_b[6], sig_c[6];
    int lds_max; // dynamic LDS bytes a block may use on this device (host-side launch choice)
    int c_kind;  // pass C kernel (DdiPassC), chosen by timing when the plan is built
    int c_tk;    // column tile of DDI_C_P2_LDS
    int rev;     // DDI_REV_* bits: passes that visit their blocks in reverse order
};

// Passes whose blocks run in reverse order, so that they first read what the pass before
// them wrote last (still held in the memory-side cache).
enum DdiReverse
{
    DDI_REV_A_FWD = 1,
    DDI_REV_B_FWD = 2,
    DDI_REV_B_INV = 4,
    DDI_REV_A_INV = 8,
};

enum DdiPassC
{
    DDI_C_GENERIC = 0, // mixed-radix LDS kernel (any length)
    DDI_C_P2_LDS  = 1, // power of two, all 3*nb lines of c_tk columns held in LDS
    DDI_C_P2_REG  = 2, // power of two, one basis atom, real spectrum: components held in registers
    DDI_C_P4_REG  = 3, // as DDI_C_P2_REG with radix-4 transforms: 4 values per lane, more lanes per CU
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

// Symmetry measures of a full kernel spectrum (filled by ddi_launch_kernel_spectrum).
struct DhatSymmetry
{
    double max_abs, max_imag, max_asym_b, max_asym_c;
};

int fft_lines_per_block( const FftLen & L );

// Pass C kernels that can run this plan, as (c_kind, c_tk) pairs (at least one).
std::vector<std::pair<int, int>> ddi_pass_c_candidates( const DdiDev & d );

// nullptr if every pass of the plan fits the device's LDS, else the reason it does not.
const char * ddi_plan_problem( const DdiDev & d );

// h = mu_i * (D * (mu s))_i / M for the spins in `spins` (SoA, nos), written to h (SoA).
// evc0/evc1 (nullable) are recorded around the convolution pass (pass C) for profiling.
void ddi_launch_field( const DdiDev & d, const double * spins, double * h, int nos, hipStream_t st,
                       hipEvent_t evc0 = nullptr, hipEvent_t evc1 = nullptr );

// Set-up of the full complex kernel spectrum Dfull[kb][ka][kc][6*n_inter] (Pb x Ha x Pc).
// `work` holds 6*n_inter*M doubles, `work2` 6*n_inter*Pc*Pb*Hs complex values.
void ddi_launch_kernel_spectrum( const DdiDev & d, const DipoleArgs & g, double * work, double2 * work2,
                                 double2 * Dfull, hipStream_t st );
// Measures realness and the kb / kc mirror symmetries of Dfull (into a device DhatSymmetry).
void ddi_launch_symmetry_check( const DdiDev & d, const double2 * Dfull, DhatSymmetry * out, hipStream_t st );
// Writes the compressed spectrum in the format described by d (dh_real, sym_b, sym_c, KB, KC).
void ddi_launch_compress( const DdiDev & d, const double2 * Dfull, void * out, hipStream_t st );

} // namespace spirit_hip
