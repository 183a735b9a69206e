// Direct-minimisation solvers of the LLG method on the GPU: VP, VP_OSO, LBFGS_OSO and
// LBFGS_Atlas, one image.
//
// Restated reference code (paths relative to the reference checkout):
//   Method_Solver<VP>::Iteration            core/include/engine/Solver_VP.hpp:29-114
//   Method_Solver<VP_OSO>::Iteration        core/include/engine/Solver_VP_OSO.hpp:33-110
//   Method_Solver<LBFGS_OSO>::Iteration     core/include/engine/Solver_LBFGS_OSO.hpp:37-72
//   Method_Solver<LBFGS_Atlas>::Iteration   core/include/engine/Solver_LBFGS_Atlas.hpp:44-101
//   Solver_Kernels (oso_*, atlas_*, maximum_rotation, lbfgs_get_searchdir)
//                                           core/src/engine/Solver_Kernels.cpp:50-244,
//                                           core/include/engine/Solver_Kernels.hpp:44-191
#pragma once

#include "llg.hpp"

namespace spirit_hip
{

constexpr int MIN_MEMORY = 3; // L-BFGS memory (n_lbfgs_memory, Solver_LBFGS_OSO.hpp:15)
constexpr int MIN_SLOTS  = 4; // partial sums per block of one reduction

// Scalars of a minimisation run, kept in device memory so that an iteration is a fixed
// sequence of launches (no host round trip; graph-capturable).
struct MinScalars
{
    double rho[MIN_MEMORY];
    double alpha[MIN_MEMORY];
    double ratio;      // VP: projection / |g|^2
    double inv_rhody2; // L-BFGS: scale of the direction after the first loop
    double coef;       // L-BFGS second loop: alpha_c - rho_c (dg_c . sd)
    double scaling;    // L-BFGS: scale of the step to the maximum move
    int vp_zero;       // VP: the projection was <= 0 (velocities reset)
    int local_iter;    // L-BFGS iteration counter
    int gd;            // this iteration takes the steepest-descent branch (first or restart)
    int flip;          // atlas: a spin left its chart (s_z a3 < -0.6)
};

// Device buffers of a minimisation run (per image SoA: component q of spin i of image k at
// [(k * dim + q) * nos + i], dim = 3, or 2 for the atlas vectors).
struct MinArgs
{
    double * v;       // VP, VP_OSO: velocities
    double * prev;    // VP: force of the previous iteration; VP_OSO: its gradient
    double * f;       // force F = -g of the current iteration
    double * grad;    // L-BFGS: OSO gradient or atlas residual
    double * grad_pr; // L-BFGS: gradient of the previous iteration
    double * sd;      // search direction
    double * q;       // L-BFGS two-loop vector
    double * da[MIN_MEMORY];
    double * dg[MIN_MEMORY];
    double * a3;          // atlas: chart of each spin (+1 / -1)
    double * part;        // per-block partial sums, [MIN_SLOTS][n_blocks]
    const double * F_ext; // chains (GNEB): the forces F_total of the images; nullptr: the LLG force -g
    MinScalars * sc;
    int n_blocks;         // noi * bpi
    int noi;              // images (1 for LLG)
    int bpi;              // blocks per image
    int dim;              // 3, or 2 for the atlas
    double maxmove;
};

bool min_is_solver( int solver );

// One iteration of `solver` on image 0 of A (spins A.s, DDI field of A.s in A.h): energy
// partials into A.part_E and max-torque partials into A.part_T (the layout k_finalize
// reduces), the effective field into A.eff on observed iterations.
void min_launch_iteration( int solver, const StepArgs & A, const MinArgs & M, int nos, hipStream_t st );

// One update of `solver` (VP_OSO, LBFGS_OSO, LBFGS_Atlas) on the M.noi images of a chain from
// their forces M.F_ext (Method_GNEB's F_total): spins A.s (nos per image), parameters A.imgs,
// run-wide stop flag A.stop.  No energy, torque or effective field (the chain's hook has them).
void min_launch_chain_update( int solver, const StepArgs & A, const MinArgs & M, int nos, hipStream_t st );

} // namespace spirit_hip
