// Discrete adjoint and tangent-linear kernels of the Depondt and Heun LLG steps (FP64).
//
// One step s -> s_new of either solver is two force evaluations (FE1 at s, FE2 at the
// predictor s'), each F = -grad E, a virtual force fv = (dtg/mu)(F + alpha s x F) + xi +
// alpha s x xi (Method_LLG.cpp:132-226), and a per-spin update (Depondt: rotation of s by
// the vector fv, Solver_Depondt.hpp:29-77; Heun: normalised Euler/trapezoid sums,
// Solver_Heun.hpp:30-83).  The energy is quadratic in the spins except for the cubic
// anisotropy, so the Hessian product H v is the gradient of the field-free Hamiltonian
// applied to v (the dipolar convolution is symmetric and is its own transpose), plus the
// diagonal cubic term.  The backward pass of one step, given lam = dL/ds_new:
//   final update adjoint           -> mu_s, mu_fv1, mu_sp, mu_g2 = -dL/dF2
//   Hessian product at s'          -> mu_sp += H mu_g2;   parameter partials at s'
//   predictor adjoint + FE1 local  -> mu_s, mu_g1
//   Hessian product at s           -> mu_s += H mu_g1;    parameter partials at s
// and lam_prev = mu_s.
#pragma once

#include "device_types.hpp"

namespace spirit_hip
{

constexpr int ADJ_THREADS    = 256;
constexpr int ADJ_MAX_SHELLS = 8;
// parameter-gradient slots: J per exchange shell, D per DMI shell, B, K per basis anisotropy
constexpr int ADJ_PG_J   = 0;
constexpr int ADJ_PG_D   = ADJ_MAX_SHELLS;
constexpr int ADJ_PG_B   = 2 * ADJ_MAX_SHELLS;
constexpr int ADJ_PG_K   = 2 * ADJ_MAX_SHELLS + 1;
constexpr int ADJ_NPG    = 2 * ADJ_MAX_SHELLS + 1 + MAX_BASIS;

struct AdjArgs
{
    HamDev H;          // the Hamiltonian of the run (masses, anisotropy layout, neighbours)
    HamDev Hh;         // the Hessian operator: H without Zeeman and cubic terms
    LLGDev L;          // LLG parameters (dt, damping, temperature, seed)
    int nos;
    int solver;        // SOLVER_DEPONDT or SOLVER_HEUN
    uint64_t step;     // index of the step (thermal-noise counter)
    double site_n[3];  // axis of the per-site anisotropy K_i
    // tape of the step
    const double * s;
    const double * F1;
    const double * fv1;
    const double * F2;
    const double * fv2;
    double * sp;       // predictor s' (recomputed)
    const double * lam;
    double * mu_s;
    double * mu_sp;
    double * mu_fv1;
    double * mu_g;     // the vector the next Hessian product acts on
    const double * hv; // dipolar field of mu_g (nullable)
    const double * xg; // configuration of the force evaluation whose Hessian / parameter terms are formed
    double * acc;      // receives += H mu_g
    double * dK;       // per-site dL/dK_i (accumulated)
    double * pg;       // [n_blocks][ADJ_NPG] parameter partials (accumulated)
};

// tangent-linear step: (s, ds) -> (s_new, ds_new) for parameter direction Hd
struct TlArgs
{
    HamDev H, Hh, Hd;  // full, Hessian, and parameter-direction (dJ, dD, dB, dK; no DDI) Hamiltonians
    LLGDev L;
    int nos;
    int solver;
    uint64_t step;
    const double * s;   // s_n
    const double * ds;  // ds_n
    double * sp;        // s'
    double * dsp;       // ds'
    double * fv1;
    double * dfv1;
    const double * h;   // dipolar field of the stage's configuration
    const double * hd;  // dipolar field of the stage's tangent
    double * s_out;
    double * ds_out;
};

void adj_launch_predictor( const AdjArgs & A, hipStream_t st );
void adj_launch_final( const AdjArgs & A, hipStream_t st );
void adj_launch_hessian( const AdjArgs & A, hipStream_t st );
void adj_launch_predictor_back( const AdjArgs & A, hipStream_t st );
void tl_launch_stage( int stage, const TlArgs & T, hipStream_t st );
int adj_n_blocks( int nos );
// one field of n vectors between AoS and SoA layouts, on the device
void device_aos_to_soa( const double * aos, double * soa, long long n, hipStream_t st );
void device_soa_to_aos( const double * soa, double * aos, long long n, hipStream_t st );

} // namespace spirit_hip
