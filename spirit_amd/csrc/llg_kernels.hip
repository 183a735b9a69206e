// Fused effective-field + LLG solver stage kernels (FP64), hand-written for gfx950.
//
// One thread owns one spin.  Each stage kernel evaluates the full local
// gradient (anisotropy, exchange, DMI, the precomputed DDI field, cubic
// anisotropy, Zeeman) of its input configuration, the LLG virtual force with
// the thermal field, and the solver's update, in registers; only the fields the
// next stage needs are written back.  Images of a batched launch (ensembles,
// GNEB chains) are addressed through a block -> image table, so one launch
// advances every image.
//
// Restated reference code (paths relative to the reference checkout):
//   Hamiltonian_Heisenberg::Gradient_and_Energy  core/src/engine/Hamiltonian_Heisenberg.cpp:704-864
//   Vectormath::idx_from_pair                    core/include/engine/Vectormath.hpp:437-528
//   Method_LLG::Calculate_Force(_Virtual)        core/src/engine/Method_LLG.cpp:113-226
//   Method_LLG::Prepare_Thermal_Field            core/src/engine/Method_LLG.cpp:66-110
//   Solver iterations                            core/include/engine/Solver_{Depondt,Heun,SIB,RK4}.hpp
//   Hook_Post_Iteration / MaxTorque_on_Image     core/src/engine/Method_LLG.cpp:247-301,
//                                                core/include/engine/Method_Solver.hpp:224-230
#include "llg_device.hpp"

namespace spirit_hip
{
// ---------------------------------------------------------------------------
// the stage kernel
// ---------------------------------------------------------------------------
// STT_G: some image takes the spin-transfer torque from the spin's derivative along the
// current (a separate instantiation: its six neighbour loads cost registers the others need)
template<int SOLVER, int STAGE, bool STT_G>
__global__ void __launch_bounds__( LLG_THREADS ) k_llg_stage( StepArgs A )
{
    if( *A.stop )
        return;
    const int blk           = xcd_contiguous_block();
    const int img           = A.blk_img[blk];
    if( A.img_stop[img] )
        return;
    const ImageDev * __restrict__ I = A.imgs + img;
    const HamDev & H        = I->ham;
    const int nos           = H.nos;
    const long long base    = 3 * I->off;
    const int i             = stage_block_order( H, blk - I->block_begin ) * LLG_THREADS + threadIdx.x;
    const bool act          = i < nos;
    constexpr bool LAST     = ( SOLVER == SOLVER_RK4 ) ? ( STAGE == 3 ) : ( STAGE == 1 );

    // input configuration of this force evaluation
    const double * x;
    if( STAGE == 0 )
        x = A.s + base;
    else if( SOLVER == SOLVER_RK4 && STAGE == 2 )
        x = A.sp2 + base;
    else
        x = A.sp + base;

    double e_i = 0, t_i = 0;
    if( act )
    {
        // the stage's own inputs are requested before the field evaluation needs its loads
        double s0[3], fv1[3];
        if( STAGE != 0 )
            load3( A.s + base, nos, i, s0 );
        constexpr bool FV1_IN = ( SOLVER == SOLVER_DEPONDT || SOLVER == SOLVER_HEUN ) && STAGE == 1;
        if( FV1_IN )
            load3( A.aux + base, nos, i, fv1 );
        double g[3], e;
        local_gradient( H, x, nos, i, ( A.h && H.has_ddi ) ? A.h + base : nullptr, g, e );
        e_i            = e;
        const double mu = H.mu[basis_of( H, i )];
        double si[3];
        load3( x, nos, i, si );
        const double F[3] = { -1 * g[0], -1 * g[1], -1 * g[2] };
        double xi[3]      = { 0, 0, 0 };
        if( I->llg.thermal && !I->llg.direct_minimization )
        {
            // drawn once per step and shared by the stages (Solver_Depondt.hpp:32)
            if( STAGE > 0 && A.xi )
                load3( A.xi + base, nos, i, xi );
            else
            {
                thermal_field( I->llg, mu, A.step_dev ? *A.step_dev + A.step : A.step, i,
                               A.noise ? A.noise + base : nullptr, xi );
                if( STAGE == 0 && A.xi )
                    store3( A.xi + base, nos, i, xi );
            }
        }
        double sg[3] = { 0, 0, 0 };
        if( STT_G && I->llg.stt == STT_GRADIENT && !I->llg.direct_minimization )
            stt_direction_gradient( H, I->llg, x, nos, i, sg );
        double fv[3];
        virtual_force( I->llg, SOLVER, H.inv_mu[basis_of( H, i )], si, F, xi, fv, sg );
        if( A.tape_F )
        {
            store3( A.tape_F + base, nos, i, F );
            store3( A.tape_fv + base, nos, i, fv );
        }
        // Hook_Post_Iteration's effective field is P_perp(s_new) F1: the force of the first
        // evaluation is stored here and projected onto the new spins when the field is read
        // (llg_launch_project_field), so an observed step adds 24 B per spin, not 72.
        if( STAGE == 0 && A.obs )
            store3( A.eff + base, nos, i, F );

        if( STAGE == 0 )
        {
            s0[0] = si[0], s0[1] = si[1], s0[2] = si[2];
        }
        double snew[3];

        if( SOLVER == SOLVER_DEPONDT )
        {
            if( STAGE == 0 )
            {
                store3( A.aux + base, nos, i, fv );
                double axis[3] = { 1 * fv[0], 1 * fv[1], 1 * fv[2] };
                double angle   = sqrt( dot3( fv, fv ) );
                normalize3_rcp( axis );
                double pr[3];
                rotate3_step( si, axis, angle, pr );
                store3( A.sp + base, nos, i, pr );
            }
            else
            {
                double t[3];
#pragma unroll
                for( int q = 0; q < 3; ++q )
                    t[q] = 0.5 * fv1[q];
#pragma unroll
                for( int q = 0; q < 3; ++q )
                    t[q] += 0.5 * fv[q];
                double angle = sqrt( dot3( t, t ) );
                normalize3_rcp( t );
                rotate3_step( s0, t, angle, snew );
                store3( ( A.s_out ? A.s_out : A.s ) + base, nos, i, snew );
            }
        }
        else if( SOLVER == SOLVER_HEUN )
        {
            if( STAGE == 0 )
            {
                store3( A.aux + base, nos, i, fv );
                double t[3], pr[3];
                neg_cross3( si, fv, t );
#pragma unroll
                for( int q = 0; q < 3; ++q )
                    pr[q] = 1 * si[q];
#pragma unroll
                for( int q = 0; q < 3; ++q )
                    pr[q] += 1 * t[q];
                normalize3_rcp( pr );
                store3( A.sp + base, nos, i, pr );
            }
            else
            {
                double t[3], t1[3];
                neg_cross3( s0, fv1, t );
#pragma unroll
                for( int q = 0; q < 3; ++q )
                    t[q] *= 0.5;
#pragma unroll
                for( int q = 0; q < 3; ++q )
                    t[q] += 1 * s0[q];
                neg_cross3( si, fv, t1 );
#pragma unroll
                for( int q = 0; q < 3; ++q )
                    t[q] += 0.5 * t1[q];
                normalize3_rcp( t );
                snew[0] = t[0], snew[1] = t[1], snew[2] = t[2];
                store3( ( A.s_out ? A.s_out : A.s ) + base, nos, i, snew );
            }
        }
        else if( SOLVER == SOLVER_SIB )
        {
            if( STAGE == 0 )
            {
                if( A.obs )
                    store3( A.aux + base, nos, i, fv );
                double pr[3];
                sib3( si, fv, pr );
#pragma unroll
                for( int q = 0; q < 3; ++q )
                    pr[q] += 1 * si[q];
#pragma unroll
                for( int q = 0; q < 3; ++q )
                    pr[q] *= 0.5;
                store3( A.sp + base, nos, i, pr );
            }
            else
            {
                if( A.obs )
                    load3( A.aux + base, nos, i, fv1 );
                sib3( s0, fv, snew );
                store3( ( A.s_out ? A.s_out : A.s ) + base, nos, i, snew );
            }
        }
        else if( SOLVER == SOLVER_RK4 )
        {
            double k[3];
            neg_cross3( si, fv, k );
            if( STAGE == 0 )
            {
                if( A.obs )
                    store3( A.aux2 + base, nos, i, fv );
                double acc[3], pr[3];
#pragma unroll
                for( int q = 0; q < 3; ++q )
                    acc[q] = 1 * si[q];
#pragma unroll
                for( int q = 0; q < 3; ++q )
                    acc[q] += 1.0 / 6.0 * k[q];
                store3( A.aux + base, nos, i, acc );
#pragma unroll
                for( int q = 0; q < 3; ++q )
                    pr[q] = 1 * si[q];
#pragma unroll
                for( int q = 0; q < 3; ++q )
                    pr[q] += 0.5 * k[q];
                normalize3_rcp( pr );
                store3( A.sp + base, nos, i, pr );
            }
            else if( STAGE == 1 || STAGE == 2 )
            {
                double acc[3], pr[3];
                load3( A.aux + base, nos, i, acc );
#pragma unroll
                for( int q = 0; q < 3; ++q )
                    acc[q] += 1.0 / 3.0 * k[q];
                store3( A.aux + base, nos, i, acc );
                const double cp = ( STAGE == 1 ) ? 0.5 : 1.0;
#pragma unroll
                for( int q = 0; q < 3; ++q )
                    pr[q] = 1 * s0[q];
#pragma unroll
                for( int q = 0; q < 3; ++q )
                    pr[q] += cp * k[q];
                normalize3_rcp( pr );
                store3( ( STAGE == 1 ? A.sp2 : A.sp ) + base, nos, i, pr );
            }
            else
            {
                double acc[3];
                load3( A.aux + base, nos, i, acc );
#pragma unroll
                for( int q = 0; q < 3; ++q )
                    acc[q] += 1.0 / 6.0 * k[q];
                normalize3_rcp( acc );
                snew[0] = acc[0], snew[1] = acc[1], snew[2] = acc[2];
                store3( ( A.s_out ? A.s_out : A.s ) + base, nos, i, snew );
                if( A.obs )
                    load3( A.aux2 + base, nos, i, fv1 );
            }
        }

        if( LAST && A.obs )
        {
            // MaxTorque_on_Image: |P_perp(s_new) F_v1|
            double d = dot3( fv1, snew );
            double pr[3];
#pragma unroll
            for( int q = 0; q < 3; ++q )
                pr[q] = fv1[q] - d * snew[q];
            t_i = dot3( pr, pr );
        }
    }
    if( LAST && A.obs )
        block_reduce_store( e_i, t_i, A.part_E, A.part_T, blk );
}

// The method constructor's force evaluation and Hook_Post_Iteration (Method_LLG.cpp:22-63,
// 247-301) on the current spins, without a thermal field: energy, max torque of the
// virtual force projected on s, and the force F (projected when the field is read).
__global__ void __launch_bounds__( LLG_THREADS ) k_observe( StepArgs A, int solver )
{
    const int blk           = xcd_contiguous_block();
    const ImageDev * __restrict__ I = A.imgs + A.blk_img[blk];
    const HamDev & H        = I->ham;
    const int nos           = H.nos;
    const long long base    = 3 * I->off;
    const int i             = ( blk - I->block_begin ) * LLG_THREADS + threadIdx.x;
    double e_i = 0, t_i = 0;
    if( i < nos )
    {
        const double * x = A.s + base;
        double g[3], e;
        local_gradient( H, x, nos, i, ( A.h && H.has_ddi ) ? A.h + base : nullptr, g, e );
        e_i = e;
        double si[3];
        load3( x, nos, i, si );
        const double F[3]  = { -1 * g[0], -1 * g[1], -1 * g[2] };
        const double xi[3] = { 0, 0, 0 };
        LLGDev P           = I->llg;
        P.thermal          = 0;
        double sg[3]       = { 0, 0, 0 };
        if( P.stt == STT_GRADIENT && !P.direct_minimization )
            stt_direction_gradient( H, P, x, nos, i, sg );
        double fv[3];
        virtual_force( P, solver, H.inv_mu[basis_of( H, i )], si, F, xi, fv, sg );
        double d = dot3( fv, si ), pr[3];
#pragma unroll
        for( int q = 0; q < 3; ++q )
            pr[q] = fv[q] - d * si[q];
        t_i = dot3( pr, pr );
        store3( A.eff + base, nos, i, F );
    }
    block_reduce_store( e_i, t_i, A.part_E, A.part_T, blk );
}

// Effective field of one image as Hook_Post_Iteration forms it (Method_LLG.cpp:284-285):
// out = P_perp(s) F = F - (F.s) s, from the force F the last observed step stored and that
// step's new spins s.
__global__ void __launch_bounds__( LLG_THREADS ) k_project_field( const double * __restrict__ F,
                                                                 const double * __restrict__ s,
                                                                 double * __restrict__ out, int nos )
{
    const int i = blockIdx.x * LLG_THREADS + threadIdx.x;
    if( i >= nos )
        return;
    double f[3], si[3], pr[3];
    load3( F, nos, i, f );
    load3( s, nos, i, si );
    const double d1 = dot3( f, si );
#pragma unroll
    for( int q = 0; q < 3; ++q )
        pr[q] = f[q] - d1 * si[q];
    store3( out, nos, i, pr );
}

void llg_launch_project_field( const double * F, const double * s, double * out, int nos, hipStream_t st )
{
    hipLaunchKernelGGL( k_project_field, dim3( ( nos + LLG_THREADS - 1 ) / LLG_THREADS ), dim3( LLG_THREADS ), 0, st,
                        F, s, out, nos );
}

void llg_launch_observe( const StepArgs & A, int solver, int n_blocks, hipStream_t st )
{
    hipLaunchKernelGGL( k_observe, dim3( n_blocks ), dim3( LLG_THREADS ), 0, st, A, solver );
}

// Gradient (SoA) and energy partials of a configuration, for getters and parity tests.
__global__ void __launch_bounds__( LLG_THREADS ) k_gradient( StepArgs A, const double * __restrict__ x0,
                                                            double * __restrict__ grad )
{
    const int blk           = xcd_contiguous_block();
    const ImageDev * __restrict__ I = A.imgs + A.blk_img[blk];
    const HamDev & H        = I->ham;
    const int nos           = H.nos;
    const long long base    = 3 * I->off;
    const int i             = ( blk - I->block_begin ) * LLG_THREADS + threadIdx.x;
    double e_i              = 0;
    if( i < nos )
    {
        double g[3], e;
        local_gradient( H, x0 + base, nos, i, ( A.h && H.has_ddi ) ? A.h + base : nullptr, g, e );
        e_i = e;
        store3( grad + base, nos, i, g );
    }
    block_reduce_store( e_i, 0.0, A.part_E, A.part_T, blk );
}

// Energy contributions per spin (SoA [term][nos] per image at 6*off), for
// Energy_Contributions_per_Spin / System_Update_Data.
__global__ void __launch_bounds__( LLG_THREADS ) k_energy_terms( StepArgs A, const double * __restrict__ x0,
                                                                double * __restrict__ out )
{
    const int blk           = xcd_contiguous_block();
    const ImageDev * __restrict__ I = A.imgs + A.blk_img[blk];
    const HamDev & H        = I->ham;
    const int nos           = H.nos;
    const long long base    = 3 * I->off;
    const int i             = ( blk - I->block_begin ) * LLG_THREADS + threadIdx.x;
    if( i < nos )
    {
        double E[N_ENERGY_TERMS];
        local_energy_terms( H, x0 + base, nos, i, ( A.h && H.has_ddi ) ? A.h + base : nullptr, E );
        for( int t = 0; t < N_ENERGY_TERMS; ++t )
            out[N_ENERGY_TERMS * I->off + (long long)t * nos + i] = E[t];
    }
}

static_assert( N_ENERGY_TERMS == N_ENERGY_TERMS_HOST, "energy term count" );

void llg_launch_energy_terms( const StepArgs & A, const double * x, double * out, int n_blocks, hipStream_t st )
{
    hipLaunchKernelGGL( k_energy_terms, dim3( n_blocks ), dim3( LLG_THREADS ), 0, st, A, x, out );
}

// Per-image reduction of the partials (one block per image, fixed order); convergence of
// each image (Method_Solver.hpp:232-239) with the index of the step at which it converged.
// Once every image has stopped, the run-wide flag stops all further stages.
constexpr int FIN_THREADS = 1024;
__global__ void __launch_bounds__( FIN_THREADS )
    k_finalize( const ImageDev * __restrict__ imgs, int n_images, const double * __restrict__ part_E,
                const double * __restrict__ part_T, double * __restrict__ obs, int * __restrict__ stop,
                int * __restrict__ img_stop, int * __restrict__ n_stopped, long long * __restrict__ stop_step,
                long long step_index, const uint64_t * __restrict__ step_dev, double force_convergence,
                int check_convergence )
{
    __shared__ double sE[FIN_THREADS], sT[FIN_THREADS];
    const int img = blockIdx.x;
    if( *stop || img_stop[img] )
        return;
    const ImageDev & I = imgs[img];
    double e = 0, t = 0;
    // thread k adds the partials k, k + blockDim.x, ... in that order; eight of them are
    // loaded before the first is added, so their latencies overlap
    constexpr int U = 8;
    const int nb    = I.n_blocks;
    for( int k0 = threadIdx.x; k0 < nb; k0 += U * blockDim.x )
    {
        double pe[U], pt[U];
#pragma unroll
        for( int u = 0; u < U; ++u )
        {
            const int k = k0 + u * blockDim.x;
            pe[u]       = k < nb ? part_E[I.block_begin + k] : 0.0;
            pt[u]       = k < nb ? part_T[I.block_begin + k] : 0.0;
        }
#pragma unroll
        for( int u = 0; u < U; ++u )
            if( k0 + u * blockDim.x < nb )
            {
                e += pe[u];
                t = fmax( t, pt[u] );
            }
    }
    sE[threadIdx.x] = e;
    sT[threadIdx.x] = t;
    __syncthreads();
    for( int w = blockDim.x / 2; w > 0; w >>= 1 )
    {
        if( threadIdx.x < w )
        {
            sE[threadIdx.x] += sE[threadIdx.x + w];
            sT[threadIdx.x] = fmax( sT[threadIdx.x], sT[threadIdx.x + w] );
        }
        __syncthreads();
    }
    if( threadIdx.x == 0 )
    {
        const double torque = sqrt( sT[0] );
        obs[2 * img]        = sE[0];
        obs[2 * img + 1]    = torque;
        if( check_convergence && torque < I.llg.force_convergence )
        {
            img_stop[img]  = 1;
            stop_step[img] = step_dev ? (long long)*step_dev + step_index : step_index;
            if( atomicAdd( n_stopped, 1 ) + 1 == n_images )
                *stop = 1;
        }
    }
}

// ---------------------------------------------------------------------------
// host launchers
// ---------------------------------------------------------------------------
template<int SOLVER, int STAGE>
static void launch_stage( const StepArgs & A, int n_blocks, hipStream_t st )
{
    if( A.stt_gradient )
        hipLaunchKernelGGL( ( k_llg_stage<SOLVER, STAGE, true> ), dim3( n_blocks ), dim3( LLG_THREADS ), 0, st, A );
    else
        hipLaunchKernelGGL( ( k_llg_stage<SOLVER, STAGE, false> ), dim3( n_blocks ), dim3( LLG_THREADS ), 0, st, A );
}

__global__ void k_step_advance( uint64_t * step, int by )
{
    if( threadIdx.x == 0 )
        *step += by;
}

void llg_launch_step_advance( uint64_t * step, int by, hipStream_t st )
{
    hipLaunchKernelGGL( k_step_advance, dim3( 1 ), dim3( 64 ), 0, st, step, by );
}

void llg_launch_stage( int solver, int stage, const StepArgs & A, int n_blocks, hipStream_t st )
{
    switch( solver )
    {
        case SOLVER_DEPONDT:
            if( stage == 0 )
                launch_stage<SOLVER_DEPONDT, 0>( A, n_blocks, st );
            else
                launch_stage<SOLVER_DEPONDT, 1>( A, n_blocks, st );
            break;
        case SOLVER_HEUN:
            if( stage == 0 )
                launch_stage<SOLVER_HEUN, 0>( A, n_blocks, st );
            else
                launch_stage<SOLVER_HEUN, 1>( A, n_blocks, st );
            break;
        case SOLVER_SIB:
            if( stage == 0 )
                launch_stage<SOLVER_SIB, 0>( A, n_blocks, st );
            else
                launch_stage<SOLVER_SIB, 1>( A, n_blocks, st );
            break;
        case SOLVER_RK4:
            if( stage == 0 )
                launch_stage<SOLVER_RK4, 0>( A, n_blocks, st );
            else if( stage == 1 )
                launch_stage<SOLVER_RK4, 1>( A, n_blocks, st );
            else if( stage == 2 )
                launch_stage<SOLVER_RK4, 2>( A, n_blocks, st );
            else
                launch_stage<SOLVER_RK4, 3>( A, n_blocks, st );
            break;
        default: break;
    }
}

void llg_launch_gradient( const StepArgs & A, const double * x, double * grad, int n_blocks, hipStream_t st )
{
    hipLaunchKernelGGL( k_gradient, dim3( n_blocks ), dim3( LLG_THREADS ), 0, st, A, x, grad );
}

void llg_launch_finalize( const StepArgs & A, int n_images, double * obs, int * img_stop, int * n_stopped,
                          long long * stop_step, long long step_index, double force_convergence, int check_convergence,
                          hipStream_t st )
{
    hipLaunchKernelGGL( k_finalize, dim3( n_images ), dim3( FIN_THREADS ), 0, st, A.imgs, n_images, A.part_E, A.part_T, obs,
                        A.stop, img_stop, n_stopped, stop_step, step_index, A.step_dev, force_convergence,
                        check_convergence );
}

} // namespace spirit_hip
