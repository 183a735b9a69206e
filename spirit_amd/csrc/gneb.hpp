// Launch interface of the GNEB chain kernels (FP64).
//
// The chain's images are stored batched, image k's SoA field at 3*k*nos, all images with
// the same number of spins.  Per-spin kernels write per-block partial sums; the last block
// of each pass to finish reduces them per image in a fixed order and runs the chain-level
// logic (reaction coordinate, tangent weights, springs, VP projection, convergence) on the
// device, so an iteration is one launch per pass and needs no host round trip.
#pragma once

#include "device_types.hpp"

namespace spirit_hip
{

constexpr int GNEB_THREADS = 256;
constexpr int GNEB_NPART   = 8;  // partial-sum slots per block
constexpr int GNEB_ISUM    = 4;  // per-image sum slots
constexpr int GNEB_COUNT_STRIDE = 32; // arrival counters 128 bytes apart
constexpr int GNEB_NCH     = 32; // scalar slots per image
constexpr int GNEB_CH_LDS  = 64; // chains up to this many images keep their scalars in LDS in the chain phases

// partial-sum slots
enum GnebPart
{
    GP_E    = 0, // energy
    GP_DIST = 1, // squared geodesic angle to the previous image
    GP_TT   = 2, // |tau_raw|^2
    GP_FT   = 3, // F_gradient . tau_raw
    GP_ET   = 4, // effective field . tau_raw
    GP_A    = 5, // general purpose (VP: v.f; path shortening; translation)
    GP_B    = 6, // general purpose (VP: f.f)
    GP_C    = 7, // general purpose (max |F|^2 of the hook; path shortening)
};

// per-image scalar slots
enum GnebChainSlot
{
    GC_E = 0,     // energy
    GC_RX,        // reaction coordinate
    GC_MODE,      // tangent rule: 0 cp*t+ + cm*t-, 1 t+, 2 t-
    GC_CP,        // weight of t+ = x_{i+1} - x_i
    GC_CM,        // weight of t- = x_i - x_{i-1}
    GC_SC,        // 1/|tau_raw|
    GC_PROJ,      // F_gradient . tau (normalised tangent)
    GC_DEDRX,     // effective field . tau
    GC_D,         // spring coefficient (normal images)
    GC_ROT,       // endpoint: rotational coefficient
    GC_TC,        // endpoint: tangential spring coefficient
    GC_PT,        // endpoint: translation projection
    GC_TORQUE,    // max torque of the image
    GC_LEN,       // path length (energy-weighted springs)
    GC_NTP,       // path shortening: 1/|x_{i+1} - x_i|
    GC_NTM,       // path shortening: 1/|x_i - x_{i-1}|
    GC_GNORM,     // path shortening: |F_gradient orthogonal|
    GC_X1,        // path shortening: first projection
    GC_X2,        // path shortening: second projection
    GC_SC2,       // path shortening: 1/|f_shrink|
    GC_SCALE,     // path shortening: scale factor
    GC_RATIO,     // image 0 only: VP projection ratio
    GC_VZERO,     // image 0 only: VP velocity reset flag
    GC_MAXT,      // image 0 only: max torque over the chain
    GC_DIST2,     // squared geodesic distance to the previous image
};

// GNEB parameters (Parameters_Method_GNEB.hpp) in device form.
struct GnebParamsDev
{
    double m; // VP mass (Method_Solver.hpp:174)
    double spring_constant;
    double spring_force_ratio;
    double path_shortening_constant;
    double delta_Rx_left, delta_Rx_right;
    double force_convergence;
    int moving_endpoints;
    int translating_endpoints;
    int escape_first;
    int check_convergence;
};

struct GnebArgs
{
    const ImageDev * imgs;
    int noi, nos, bpi, n_blocks; // images, spins per image, blocks per image, blocks
    const double * x;            // configuration of this force evaluation
    const double * xc;           // the chain's current spins (path shortening, translation)
    const double * h;            // DDI field of x (nullable)
    double * eff;                // effective field -g(x)
    double * Fg;                 // F_gradient, projected on the tangent planes of x
    double * tau;                // raw tangents
    double * Ft;                 // F_total (= the method's forces)
    double * ftr;                // translation forces of the two endpoints (image 0 / noi-1 slots)
    double * fv;                 // virtual forces (non-VP solvers)
    double * v;                  // VP velocities
    double * fprev;              // VP forces of the previous iteration
    double * part;               // [n_blocks][GNEB_NPART]
    double * ch;                 // [noi][GNEB_NCH]
    const int * types;           // image types (0 normal, 1 climbing, 2 falling, 3 stationary)
    int * stop;                  // 0 running, 1 converged, 2 two images coincide
    long long * iter;            // iterations done
    int * count;                 // arrival counters of the running pass (all 0 between passes): [0] images whose
                                 // sums are stored, [GNEB_COUNT_STRIDE*(1+img)] blocks of image img done
    double * isum;               // [noi][GNEB_ISUM] per-image sums of the running pass's partials
    GnebParamsDev P;
};

// per-spin passes
enum GnebPass
{
    GPASS_GRAD = 0,     // gradient, energy, geodesic distance
    GPASS_TANGENT,      // raw tangents, F_gradient projection, their dots
    GPASS_TRANSLATE,    // translating endpoints
    GPASS_SHRINK1,      // path shortening rounds
    GPASS_SHRINK2,
    GPASS_SHRINK3,
    GPASS_SHRINK4,
    GPASS_FORCE_VP,     // F_total, VP velocity update
    GPASS_FORCE_FV,     // F_total, virtual force
    GPASS_VP_UPDATE,    // VP configuration update + Hook_Post_Iteration
    GPASS_HOOK,         // Hook_Post_Iteration (non-VP solvers)
};

// chain phases (run by the last block of the pass before them)
enum GnebPhase
{
    GPH_DIST = 0, // energies, Rx, tangent rules
    GPH_TANGENT,  // tangent norms, projections, springs, endpoint coefficients
    GPH_TRANSLATE,
    GPH_SHRINK1,
    GPH_SHRINK2,
    GPH_SHRINK3,
    GPH_SHRINK4,
    GPH_VP,       // VP projection ratio
    GPH_HOOK,     // max torque, convergence, iteration count
};

// One per-spin pass over the chain, whose last block to finish runs the chain phase that
// consumes the pass's partial sums (GRAD -> DIST, TANGENT -> TANGENT, TRANSLATE -> TRANSLATE,
// SHRINKn -> SHRINKn, FORCE_VP -> VP, VP_UPDATE and HOOK -> HOOK; FORCE_FV has none).
void gneb_launch_pass( int pass, const GnebArgs & A, hipStream_t st );
// Solver stages of the predictor-corrector solvers on the chain (virtual forces already computed):
// stage 0 writes the predictor, stage 1 the new configuration (RK4: stages 0..3).
struct GnebSolverArgs
{
    int noi, nos, bpi;
    double * x;                 // configurations (updated by the last stage)
    double * pred;              // predictor
    double * tmp;               // Heun: -(x x fv); RK4: accumulator
    const double * fv;          // virtual force of this stage's input
    const double * fv0;         // virtual force of the first evaluation (Depondt, Heun corrector)
    const int * stop;
};
void gneb_launch_solver_stage( int solver, int stage, const GnebSolverArgs & S, hipStream_t st );

} // namespace spirit_hip
