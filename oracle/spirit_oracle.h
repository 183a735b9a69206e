/*
 * spirit_oracle.h -- CPU restatement (C99 + OpenMP, FP64) of the reference's
 * effective-field and LLG step path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: it is loaded by
 * tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg, and
 * by nothing else.  The product (spirit_amd/libSpirit.so) never links it.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the reference checkout, e.g. core/src/engine/Hamiltonian_Heisenberg.cpp).
 *
 * Spin / gradient arrays crossing this interface are AoS doubles
 * [x0,y0,z0,x1,...] exactly like the reference's vectorfield.
 */
#ifndef SPIRIT_ORACLE_H
#define SPIRIT_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MAX_BASIS 8
#define ORC_MAX_PAIRS 256
#define ORC_MAX_QUADS 64

/* Physical constants, core/include/utility/Constants.hpp:18-40 */
#define ORC_MU_B 0.057883817555
#define ORC_MU_0 ( 2.0133545 * 1e-28 )
#define ORC_K_B 0.08617330350
#define ORC_GAMMA 0.1760859644
#define ORC_PI 3.141592653589793238462643383279502884197169399375105820974

/* Solver ids, core/include/Spirit/Simulation.h:33-54 */
#define ORC_SOLVER_VP 0
#define ORC_SOLVER_SIB 1
#define ORC_SOLVER_DEPONDT 2
#define ORC_SOLVER_HEUN 3
#define ORC_SOLVER_RK4 4
#define ORC_SOLVER_LBFGS_OSO 5
#define ORC_SOLVER_LBFGS_ATLAS 6
#define ORC_SOLVER_VP_OSO 7

/* DDI methods, core/include/Spirit/Hamiltonian.h:47-57 */
#define ORC_DDI_NONE 0
#define ORC_DDI_FFT 1
#define ORC_DDI_CUTOFF 3

typedef struct orc_params
{
    /* ---- geometry: core/src/data/Geometry.cpp:22-161 ---- */
    int n_cells[3];
    int n_cell_atoms;
    double bravais[3][3];                 /* bravais[k] is the k-th Bravais vector */
    double cell_atoms[ORC_MAX_BASIS][3];  /* fractional coordinates (basis of Bravais vectors) */
    double lattice_constant;
    double mu_s[ORC_MAX_BASIS];           /* per basis atom */
    int bc[3];                            /* periodic boundary conditions a,b,c */

    /* ---- Heisenberg Hamiltonian: Hamiltonian_Heisenberg.hpp ---- */
    double field_tesla;      /* |B| in Tesla; mu_B folded in internally (Hamiltonian_Heisenberg.cpp:35) */
    double field_normal[3];
    int n_aniso;
    int aniso_idx[ORC_MAX_BASIS];
    double aniso_K[ORC_MAX_BASIS];
    double aniso_normal[ORC_MAX_BASIS][3];
    int n_cubic;
    int cubic_idx[ORC_MAX_BASIS];
    double cubic_K[ORC_MAX_BASIS];
    /* optional per-site uniaxial anisotropy K_i along site_aniso_normal (extension; NULL = off) */
    const double * site_aniso_K;
    double site_aniso_normal[3];

    /* pair tables (redundant, gather-style, as the OpenMP build uses them) */
    int n_exch;
    int exch[ORC_MAX_PAIRS][5]; /* i, j, da, db, dc */
    double exch_J[ORC_MAX_PAIRS];
    int n_dmi;
    int dmi[ORC_MAX_PAIRS][5];
    double dmi_D[ORC_MAX_PAIRS];
    double dmi_normal[ORC_MAX_PAIRS][3];

    /* dipole-dipole */
    int ddi_method;
    int ddi_n_periodic_images[3];
    int ddi_pb_zero_padding;
    double ddi_radius;

    /* quadruplets (Hamiltonian_Heisenberg.hpp, Dataparser.cpp:521-680): basis atoms i, j, k, l and
       the translations d_j, d_k, d_l from i's cell, magnitude Q */
    int n_quad;
    int quad[ORC_MAX_QUADS][13]; /* i, j, k, l, d_j[3], d_k[3], d_l[3] */
    double quad_Q[ORC_MAX_QUADS];
} orc_params;

/* ---------------- geometry / neighbour tables ---------------- */

/* Shell radii, core/src/engine/Neighbours.cpp:15-81. */
void orc_shell_radii( const orc_params * p, int n_shells, double * radii );

/* Neighbour pairs in the first n_shells shells, core/src/engine/Neighbours.cpp:83-156.
   out[k*6 + 0..5] = i, j, da, db, dc, shell.  Returns the number of pairs (or -needed if max too small). */
int orc_neighbours_in_shells( const orc_params * p, int n_shells, int redundant, int * out, int max_pairs );

/* DM normal of a pair, core/src/engine/Neighbours.cpp:249-286. */
void orc_dmi_normal( const orc_params * p, const int pair[5], int chirality, double out[3] );

/* Fill exch/dmi tables from shells, restating Hamiltonian_Heisenberg::Update_Interactions
   (core/src/engine/Hamiltonian_Heisenberg.cpp:101-160) with use_redundant_neighbours = true. */
int orc_setup_shells( orc_params * p, int n_shells_exch, const double * jij, int n_shells_dmi, const double * dij,
                      int chirality );

/* Append the reversed copies of user-given pair lists (Hamiltonian_Heisenberg.cpp:121-133, 152-163). */
void orc_make_pairs_redundant( orc_params * p );

/* Neighbour index of a pair, core/include/engine/Vectormath.hpp:437-528 (defects off). */
int orc_idx_from_pair( int ispin, const int bc[3], const int n_cells[3], int N, const int pair[5] );

/* ---------------- system ---------------- */
typedef struct orc_sys orc_sys;

orc_sys * orc_create( const orc_params * p );
void orc_destroy( orc_sys * s );
int orc_nos( const orc_sys * s );
/* positions, Geometry.cpp:142-161 */
void orc_positions( const orc_sys * s, double * out );
/* padded DDI grid, Hamiltonian_Heisenberg.cpp:1501-1530 (without the kissFFT length-1 workaround) */
void orc_ddi_padded( const orc_sys * s, int out[3] );
/* real-space dipole tensor on the padded grid, Hamiltonian_Heisenberg.cpp:1406-1499.
   out layout: [b_inter][comp 6][c][b][a] (a fastest). */
void orc_ddi_dipole_matrices( const orc_sys * s, double * out );

/* the cutoff method's pair list (Neighbours::Get_Pairs_in_Radius, Neighbours.cpp:158-246) with
   distances and unit vectors (DDI_from_Pair, :288-307); returns the number of pairs */
int orc_ddi_cutoff_pairs( const orc_sys * s, int * pairs, double * magnitudes, double * normals, int max_pairs );

/* E_Quadruplet per spin, Hamiltonian_Heisenberg.cpp:513-557 (added into energy, nos values) */
void orc_energy_quadruplet( orc_sys * s, const double * spins, double * energy );

/* Hamiltonian_Heisenberg::Hessian, Hamiltonian_Heisenberg.cpp:1119-1245: dense 3N x 3N,
   row-major, index 3*ispin + component */
void orc_hessian( orc_sys * s, const double * spins, double * H );

/* Gradient_and_Energy, Hamiltonian_Heisenberg.cpp:704-766 */
void orc_gradient_and_energy( orc_sys * s, const double * spins, double * grad, double * energy );
/* Gradient (Zeeman first), Hamiltonian_Heisenberg.cpp:670-702 */
void orc_gradient( orc_sys * s, const double * spins, double * grad );
/* Single contributions (gradient added into grad): 0 zeeman 1 aniso 2 cubic 3 exchange 4 dmi
   5 ddi (the configured method: fft, cutoff with radius >= 0, direct for a negative radius)
   6 ddi-direct (Gradient_DDI_Direct, :1016-1071) 7 quadruplets (Gradient_Quadruplet, :1073-1117) */
void orc_gradient_term( orc_sys * s, int term, const double * spins, double * grad );

/* ---------------- LLG ---------------- */
typedef struct orc_llg
{
    double dt;                 /* ps */
    double damping;
    double temperature;        /* K */
    int direct_minimization;
    uint64_t seed;             /* Philox key for the thermal field */
    const double * noise;      /* if non-NULL: injected N(0,1) values, [n_steps][nos][3] */
    /* Parameters_Method_LLG.hpp:19-38 */
    double beta;               /* non-adiabatic STT parameter */
    int stt_use_gradient;      /* 1: the spin derivative along the current, 0: monolayer */
    double stt_magnitude;      /* a_j; STT acts for a_j > 0 */
    double stt_normal[3];      /* polarisation / current direction (unit) */
    double tgrad_inclination;  /* K per length unit of the positions */
    double tgrad_direction[3];
} orc_llg;

/* Vectormath::jacobian (core/src/engine/Vectormath.cpp:816-898) of the vector field vf (AoS),
   applied to the direction je as Method_LLG.cpp:188-196 does: out_i = J_i je */
void orc_jacobian_direction( const orc_sys * s, const double * vf, const double je[3], double * out );

/* Vectormath::get_gradient_distribution (core/src/engine/Vectormath.cpp:633-652) with the
   cut to [0, 1e30] of Method_LLG.cpp:84-86: the per-site temperature */
void orc_temperature_distribution( const orc_sys * s, const double direction[3], double T, double inclination,
                                   double * out );

/* Iterate n_steps LLG steps in place (Method_LLG.cpp:66-301 + Solver_*.hpp).
   step0: global index of the first step (thermal-noise counter).
   Outputs (nullable): energy (of the last force evaluation), max torque
   (Method_Solver.hpp:224-230), effective field (Method_LLG.cpp:287-288). */
void orc_llg_iterate( orc_sys * s, int solver, const orc_llg * par, double * spins, int n_steps, uint64_t step0,
                      double * energy, double * max_torque, double * eff_field );

/* As orc_llg_iterate, with Method::Iterate's convergence test (Method.cpp:57-113,
   n_iterations_amortize = 1; Method_LLG.cpp:247-270): stops after the first iteration whose
   max torque is below force_convergence.  Returns the number of iterations done.  The
   direct-minimisation solvers VP, VP_OSO, LBFGS_OSO and LBFGS_Atlas keep their state
   (velocities, L-BFGS memory, atlas charts) across the iterations of one call, as one
   method object does; every call starts them afresh, as Simulation_LLG_Start does. */
long orc_llg_run( orc_sys * s, int solver, const orc_llg * par, double * spins, long n_iterations,
                  double force_convergence, double * energy, double * max_torque, double * eff_field );

/* Thermal N(0,1) triple for site i at step `step` (Philox4x32-10 + Box-Muller) -- the build's own
   counter-based generator that replaces the reference's sequential mt19937 stream. */
void orc_normal3( uint64_t seed, uint64_t step, uint64_t site, double out[3] );

/* ---------------- GNEB ---------------- */
typedef struct orc_gneb
{
    double dt;                      /* LLG time step of the images [ps] (Method_GNEB.cpp:379, Solver_VP.hpp:92) */
    double spring_constant;
    double spring_force_ratio;      /* 0: Rx springs; > 0: energy-weighted path lengths */
    double path_shortening_constant;
    int moving_endpoints;
    int translating_endpoints;
    int escape_first;
    double delta_Rx_left, delta_Rx_right;
    int n_E_interpolations;
    double force_convergence;
} orc_gneb;

/* Method_GNEB with solver VP / SIB / Depondt / Heun / RK4 / VP_OSO / LBFGS_OSO / LBFGS_Atlas (Method_GNEB.cpp:88-456,
   Solver_*.hpp, Method_Solver.hpp:224-239, Method.cpp:57-113 with n_iterations_amortize = 1).
   spins: noi images of nos spins (AoS), advanced in place.  image_type: 0 normal,
   1 climbing, 2 falling, 3 stationary (Spin_System_Chain.hpp:13-19).  A fresh solver state
   (zero velocities) is used, as each Simulation_GNEB_Start constructs a new method.
   Outputs (nullable): energies/Rx/max torque per image and effective fields of the last
   force evaluation; interpolated Rx/E (noi + (noi-1)*n_E_interpolations points).
   Returns the number of iterations done, or -(iterations+1) if the run stopped because
   two neighbouring images coincide. */
long orc_gneb_run( orc_sys * s, int solver, const orc_gneb * par, int noi, const int * image_type, double * spins,
                   long n_iterations, int check_convergence, double * energies, double * Rx, double * max_torque_all,
                   double * eff_field, double * Rx_interp, double * E_interp );

/* Cubic_Hermite_Spline::Interpolate, core/src/utility/Cubic_Hermite_Spline.cpp:10-47 */
void orc_cubic_hermite( const double * x, const double * p, const double * m, int n, int n_interp, double * out_x,
                        double * out_p );

/* Configuration_Chain::Homogeneous_Rotation, core/src/utility/Configuration_Chain.cpp:28-75 */
void orc_homogeneous_rotation( double * chain, int nos, int idx_1, int idx_2 );

/* ---------------- FFT (used by the DDI convolution) ---------------- */
/* In-place complex DFT of length n (interleaved re,im), sign -1 forward, +1 backward, unnormalised. */
void orc_fft( double * data, int n, int sign );
/* 1: the lines of each 3-D transform run on all OpenMP threads; 0: on one thread (the
   reference's kissFFT build transforms serially) */
void orc_set_fft_parallel( int on );

/* ---------------- utilities ---------------- */
/* Rodrigues rotation, core/src/engine/Vectormath.cpp:474-479 */
void orc_rotate( const double v[3], const double axis[3], double angle, double out[3] );
/* sib_transform, core/src/engine/Solver_Kernels.cpp:15-48 (one spin) */
void orc_sib_transform( const double s[3], const double f[3], double out[3] );

#ifdef __cplusplus
}
#endif

#endif
