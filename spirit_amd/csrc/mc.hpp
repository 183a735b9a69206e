// Metropolis Monte Carlo of one image on the GPU (Method_MC, core/src/engine/Method_MC.cpp:40-200).
#pragma once

#include "device_types.hpp"

namespace spirit_hip
{

// Arguments of the Monte Carlo kernels.  Spins of one colour class share no exchange or DMI
// pair, so a class is updated in parallel, each trial against the current neighbours.
struct McArgs
{
    const ImageDev * img;
    double * s;            // SoA spins of the image
    int nos;
    int K;                 // colour classes: colour(a, b, c) = (ca a + cb b + cc c) mod K
    int ca, cb, cc;
    int step_cone;         // trial moves inside a cone around the spin (else the reference's sphere draw)
    int adaptive;          // adapt the cone to the target acceptance ratio
    uint64_t seed;
    double temperature;    // K
    double kT;             // k_B T [meV]
    double target;         // target acceptance ratio
    double * cone;         // cone angle [rad], adapted on the device, and its cosine
    double * part;         // per-block rejection counts, [K][n_blocks]
    int n_blocks;          // blocks of the one-thread-per-site sweep (part's row length)
    int inv_ca;            // inverse of ca mod K with one basis atom, else 0: per-colour sweeps
    int groups;            // ceil(Na / K): groups of K cells per row
    int n_blocks_rows;     // blocks of the per-colour sweep (<= n_blocks)
    const uint64_t * step; // Monte Carlo step index (random-number counter)
};

// One Metropolis iteration: the cone feedback from the previous iteration's rejections, then
// one sweep over every colour class; advances *step by one.
void mc_launch_iteration( const McArgs & A, uint64_t * step, hipStream_t st );

} // namespace spirit_hip
