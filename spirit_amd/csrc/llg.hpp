// Launch interface of the fused effective-field + LLG stage kernels.
#pragma once

#include "device_types.hpp"

namespace spirit_hip
{

constexpr int LLG_THREADS = 256;

// Arguments of one stage launch.  Fields are batched SoA buffers; image k's
// three components start at 3*imgs[k].off.
struct StepArgs
{
    const ImageDev * imgs; // device array of images
    const int * blk_img;   // device array: block -> image
    double * s;            // current spins (updated by the last stage)
    double * sp;           // predictor spins
    double * sp2;          // second predictor buffer (RK4)
    double * aux;          // first virtual force (Depondt/Heun, SIB on observed steps) or RK4 accumulator
    double * aux2;         // RK4: first virtual force on observed steps
    const double * h;      // DDI field of the stage's input configuration (nullable)
    double * eff;          // observed steps: the first evaluation's force F1 (LLG stages) or the
                           // projected effective field (direct-minimisation kernels)
    const double * noise;  // injected N(0,1) thermal noise of this step, AoS per spin (nullable)
    double * xi;           // the step's thermal field: drawn and stored by the first stage, read by
                           // the later ones (nullable: every stage draws it)
    double * part_E;       // per-block energy partial sums
    double * part_T;       // per-block max squared torque
    int * stop;            // set once every image has stopped; every stage returns early
    const int * img_stop;  // per image: set once that image has stopped (converged or ended)
    double * s_out;        // where the last stage writes the new spins (nullptr: in place into s)
    double * tape_F;       // this stage's force F = -g (nullable; adjoint tape)
    double * tape_fv;      // this stage's virtual force (nullable; adjoint tape)
    uint64_t step;         // global step index (thermal-noise counter); with step_dev, the offset from *step_dev
    const uint64_t * step_dev; // base step index in device memory (graph replays); nullptr: step is absolute
    int obs;               // compute observables (energy, torque, effective field)
    int stt_gradient;      // some image's STT uses the spin derivative along the current (STT_GRADIENT)
};

void llg_launch_stage( int solver, int stage, const StepArgs & A, int n_blocks, hipStream_t st );
// *step += by on the device (the step counter of graph-replayed step sequences)
void llg_launch_step_advance( uint64_t * step, int by, hipStream_t st );
void llg_launch_gradient( const StepArgs & A, const double * x, double * grad, int n_blocks, hipStream_t st );
// per-spin energy contributions (Zeeman, anisotropy, cubic, exchange, DMI, DDI), SoA [6][nos] per
// image at 6*off; the DDI field of x must be in A.h
// per-spin energy contributions written by llg_launch_energy_terms: Zeeman, anisotropy, cubic
// anisotropy, exchange, DMI, DDI, quadruplets ([term][spin] per image)
constexpr int N_ENERGY_TERMS_HOST = 7;
void llg_launch_energy_terms( const StepArgs & A, const double * x, double * out, int n_blocks, hipStream_t st );
void llg_launch_observe( const StepArgs & A, int solver, int n_blocks, hipStream_t st );
// out = F - (F.s) s per spin (one image's SoA fields)
void llg_launch_project_field( const double * F, const double * s, double * out, int nos, hipStream_t st );
// Per-image energy / max torque into obs[2*img], [2*img+1]; with check_convergence an image whose
// torque is below force_convergence stops (img_stop, stop_step[img]); the run-wide stop flag is set
// when n_stopped reaches n_images.
void llg_launch_finalize( const StepArgs & A, int n_images, double * obs, int * img_stop, int * n_stopped,
                          long long * stop_step, long long step_index, double force_convergence, int check_convergence,
                          hipStream_t st );

} // namespace spirit_hip
