// Plain-old-data structures shared between the host engine and the HIP kernels.
// Spin-like fields live in HBM as structure-of-arrays per image:
//   field[q * nos + ispin], q = 0,1,2 (x,y,z), ispin = ib + nb*(a + Na*(b + Nb*c))
// (the reference's index convention, core/src/data/Geometry.cpp:150-157).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace spirit_hip
{

constexpr int MAX_BASIS      = 8;
constexpr int MAX_PAIRS      = 64;
constexpr int MAX_FFT_STAGES = 16;
constexpr int GATHER_PAIRS   = 6; // pair tables up to this size are gathered into registers per spin

// Physical constants, core/include/utility/Constants.hpp:18-40
constexpr double mu_B  = 0.057883817555;
constexpr double mu_0  = 2.0133545 * 1e-28;
constexpr double k_B   = 0.08617330350;
constexpr double gamma_e = 0.1760859644; // gyromagnetic ratio of the electron [rad/(ps*T)]
constexpr double Pi    = 3.141592653589793238462643383279502884197169399375105820974;

enum SolverId
{
    SOLVER_VP          = 0,
    SOLVER_SIB         = 1,
    SOLVER_DEPONDT     = 2,
    SOLVER_HEUN        = 3,
    SOLVER_RK4         = 4,
    SOLVER_LBFGS_OSO   = 5,
    SOLVER_LBFGS_ATLAS = 6,
    SOLVER_VP_OSO      = 7,
};

// Division by an invariant positive integer d for numerators 0 <= n < 2^31:
// floor(n / d) = (n * m) >> shift with shift = 31 + ceil(log2 d) and m = floor(2^shift / d) + 1
// (m < 2^33, held as m_lo + 2^32 m_hi).  The rounding error of m stays below 1/d for such n.
struct FastDiv
{
    uint32_t m_lo, m_hi;
    int shift, d;
};

// A neighbour pair (i, j, translation) with its exchange constant J, or its DM
// magnitude D and unit DM vector n (Vectormath_Defines.hpp:98-107).
struct PairDev
{
    int i, j, da, db, dc, pad;
    int delta;  // index offset j - i of the neighbour for a site away from every boundary
    int pad2;
    double J;
    double D;
    double n[3];
};

// A quadruplet (Vectormath_Defines.hpp:82-86) with its magnitude Q.
struct QuadDev
{
    int i, j, k, l;
    int dj[3], dk[3], dl[3];
    int pad;
    double Q;
};

// Everything the fused local-field kernel needs for one image.
struct HamDev
{
    int Na, Nb, Nc, nb; // cells and basis atoms
    int nos;
    int bc[3];
    int n_exch, n_dmi, n_aniso, n_cubic;
    int has_zeeman;
    int has_ddi;
    int has_site_K;
    int dmi_same;            // DMI pair k has exchange pair k's (i, j, translation) for every k < n_dmi
    int halo[3];             // largest |translation| of an exchange pair along a, b, c
    int row_tile_log2;       // stage kernels sweep tiles of 2^row_tile_log2 rows along b through the c planes
    FastDiv div_nb, div_Na, div_Nb; // index -> (basis atom, a, b, c) without hardware division
    double B;                // external_field_magnitude * mu_B (Hamiltonian_Heisenberg.cpp:35)
    double Bn[3];            // external_field_normal
    double mu[MAX_BASIS];    // mu_s per basis atom
    double inv_mu[MAX_BASIS]; // 1 / mu_s
    int aniso_idx[MAX_BASIS];
    double aniso_K[MAX_BASIS];
    double aniso_n[MAX_BASIS][3];
    int cubic_idx[MAX_BASIS];
    double cubic_K[MAX_BASIS];
    double site_n[3];
    const double * site_K;   // per-site uniaxial K_i (nullable)
    const PairDev * exch;    // n_exch redundant exchange pairs
    const PairDev * dmi;     // n_dmi redundant DMI pairs
    int n_quad;              // quadruplets
    int quad_in_step;        // Gradient_and_Energy adds them (Hamiltonian_Heisenberg.cpp:754)
    const QuadDev * quad;
    // the first GATHER_PAIRS exchange and DMI pairs, held in the record itself: the gathered
    // field evaluation reads them with the record (scalar loads), not through a pointer
    PairDev exch_g[GATHER_PAIRS];
    PairDev dmi_g[GATHER_PAIRS];
};

// Spin-transfer torque of the dynamics' virtual force (Method_LLG.cpp:182-212).
enum SttMode
{
    STT_OFF       = 0,
    STT_MONOLAYER = 1, // fixed polarisation p
    STT_GRADIENT  = 2, // p = (j . grad) s, the spin's derivative along the current direction
};

// LLG parameters of one image (Method_LLG.cpp:132-226, 66-110).
struct LLGDev
{
    double dt;
    double damping;
    double temperature;
    int direct_minimization;
    int thermal; // temperature > 0 or a temperature gradient
    uint64_t seed;
    double force_convergence; // the image stops once its max torque is below this
    double vf_dtg;            // dt gamma / mu_B / (1 + damping^2), the dynamics' virtual-force scale
    // STT: the virtual force gains stt_c1 p + stt_c2 (p x s) after the division by mu_s
    int stt;                  // SttMode
    double stt_c1, stt_c2;
    double stt_p[3];          // monolayer polarisation
    double stt_w[3];          // gradient: B^-1 j, the weights of the differences along the three
                              // lattice translations (B: columns lattice_constant * bravais_k)
    const double * temp_site; // per-site temperature of a temperature gradient (nullable: homogeneous)
};

// One image of a batched launch: where its fields live and how it is parameterised.
struct ImageDev
{
    HamDev ham;
    LLGDev llg;
    long long off;      // spin offset of this image inside the batched SoA buffers (fields at 3*off)
    int block_begin;    // first block of this image in the launch
    int n_blocks;       // blocks of this image
};

// One FFT length: Stockham stage radices and the n-th roots of unity exp(-2 pi i e / n).
// For powers of two >= 8, tw2 holds the twiddles of the register-resident radix-8
// plan in stage order: stage s (>= 1, radix R, Ns = 8^s) at offset 8^s - 8,
// entry jm*(R-1) + r-1 = w_n^(r * jm * n/(Ns*R)).
struct FftLen
{
    int n;
    int n_stages;
    int radix[MAX_FFT_STAGES];
    const double2 * W;
    const double2 * tw2;
};

} // namespace spirit_hip
