// Direct-minimisation solvers (VP, VP_OSO, LBFGS_OSO, LBFGS_Atlas) of the LLG method, FP64,
// hand-written for gfx950.
//
// An iteration is a fixed sequence of launches: per-spin kernels (one thread per spin) and,
// between them, a one-block kernel that reduces the per-block partial sums in a fixed order
// and applies the solver's scalar logic (velocity projection, L-BFGS coefficients, step
// scaling, restarts, atlas chart changes) to a state kept in device memory.  The host issues
// the same launches every iteration, so runs need no host round trip and replay as graphs.
//
// Restated reference code (paths relative to the reference checkout):
//   Solver_VP.hpp:29-114, Solver_VP_OSO.hpp:33-110, Solver_LBFGS_OSO.hpp:37-72,
//   Solver_LBFGS_Atlas.hpp:44-101 (core/include/engine/);
//   Solver_Kernels::oso_calc_gradients, oso_rotate, maximum_rotation, atlas_rotate,
//   atlas_calc_gradients, ncg_atlas_check_coordinates, lbfgs_atlas_transform_direction
//                                      core/src/engine/Solver_Kernels.cpp:50-244
//   Solver_Kernels::lbfgs_get_searchdir core/include/engine/Solver_Kernels.hpp:51-191
//   Method_LLG::Calculate_Force(_Virtual), Hook_Post_Iteration
//                                      core/src/engine/Method_LLG.cpp:113-301
#include "llg_device.hpp"
#include "minimize.hpp"

namespace spirit_hip
{

enum MinPhase
{
    PH_VP = 0,  // VP: ratio = (v . g) / (g . g)
    PH_RHO,     // L-BFGS: rho of the newest pair, first alpha, scale after the first loop; restarts
    PH_ALPHA,   // L-BFGS first loop: alpha_c = rho_c (da_c . q)
    PH_COEF,    // L-BFGS second loop: alpha_c - rho_c (dg_c . sd)
    PH_SCALE,   // L-BFGS: step scaling to the maximum move; iteration counter
    PH_FLIP,    // atlas: did a spin leave its chart
    PH_FLIPRHO, // atlas: rho after the chart change
};

constexpr double LBFGS_EPS = 1e-300; // Solver_Kernels.hpp:59 (double)
constexpr double MIN_MASS  = 1.0;    // Method_Solver.hpp:174
constexpr int SC_THREADS   = 1024;

// LLG runs stop on the run-wide flag or image 0's; chains (forces given) on the chain's flag
__device__ __forceinline__ bool min_stopped( const StepArgs & A, const MinArgs & M )
{
    return *A.stop || ( !M.F_ext && A.img_stop[0] );
}

// the spin of this thread: image img (blocks img*bpi ..), spin i of the image, and the
// offsets of the image's 3-vector, dim-vector and scalar fields
struct MinSite
{
    int img, i;
    long long b3, bd, b1;
};
__device__ __forceinline__ MinSite min_site( int bpi, int dim, int nos )
{
    MinSite m;
    m.img = blockIdx.x / bpi;
    m.i   = ( blockIdx.x - m.img * bpi ) * LLG_THREADS + threadIdx.x;
    m.b3  = 3LL * m.img * nos;
    m.bd  = (long long)dim * m.img * nos;
    m.b1  = (long long)m.img * nos;
    return m;
}

// first loop of the two-loop recursion, k = M-1 .. 0
__device__ __forceinline__ int lbfgs_c1( int li, int k )
{
    return ( k + li % MIN_MEMORY + 1 ) % MIN_MEMORY;
}
// second loop, k = 0 .. M-1
__device__ __forceinline__ int lbfgs_c2( int li, int k )
{
    return li < MIN_MEMORY ? k : ( k + li % MIN_MEMORY + 1 ) % MIN_MEMORY;
}

// fixed-order block sums of NS values; thread 0 gets the totals in x
template<int NS>
__device__ __forceinline__ void min_block_sums( double ( &x )[NS] )
{
    __shared__ double red[NS][LLG_THREADS / 64];
#pragma unroll
    for( int s = 0; s < NS; ++s )
    {
#pragma unroll
        for( int off = 32; off > 0; off >>= 1 )
            x[s] += __shfl_down( x[s], off, 64 );
    }
    const int w = threadIdx.x / 64, lane = threadIdx.x % 64;
    if( lane == 0 )
    {
#pragma unroll
        for( int s = 0; s < NS; ++s )
            red[s][w] = x[s];
    }
    __syncthreads();
    if( threadIdx.x == 0 )
    {
#pragma unroll
        for( int s = 0; s < NS; ++s )
        {
            double t = red[s][0];
            for( int k = 1; k < LLG_THREADS / 64; ++k )
                t += red[s][k];
            x[s] = t;
        }
    }
}

__device__ __forceinline__ double min_block_max( double t )
{
    __shared__ double red[LLG_THREADS / 64];
#pragma unroll
    for( int off = 32; off > 0; off >>= 1 )
        t = fmax( t, __shfl_down( t, off, 64 ) );
    const int w = threadIdx.x / 64, lane = threadIdx.x % 64;
    if( lane == 0 )
        red[w] = t;
    __syncthreads();
    if( threadIdx.x == 0 )
    {
        t = red[0];
        for( int k = 1; k < LLG_THREADS / 64; ++k )
            t = fmax( t, red[k] );
    }
    return t;
}

// Solver_Kernels::oso_calc_gradients: t (-s x F), t = ((0,0,1),(0,-1,0),(1,0,0))
__device__ __forceinline__ void oso_gradient( const double * s, const double * F, double * g )
{
    double c[3];
    cross3( s, F, c );
    const double w[3] = { -c[0], -c[1], -c[2] };
    g[0]              = w[2];
    g[1]              = -w[1];
    g[2]              = w[0];
}

// Solver_Kernels::atlas_calc_gradients (one spin)
__device__ __forceinline__ void atlas_gradient( const double * s, const double * f, double a3, double * g )
{
    const double J00 = s[1] * s[1] + s[2] * ( s[2] + a3 );
    const double J10 = -s[0] * s[1];
    const double J01 = -s[0] * s[1];
    const double J11 = s[0] * s[0] + s[2] * ( s[2] + a3 );
    const double J02 = -s[0] * ( s[2] + a3 );
    const double J12 = -s[1] * ( s[2] + a3 );
    g[0]             = -( J00 * f[0] + J01 * f[1] + J02 * f[2] );
    g[1]             = -( J10 * f[0] + J11 * f[1] + J12 * f[2] );
}

// Solver_Kernels::oso_rotate (one spin): rotation by exp of the skew matrix of sd
__device__ __forceinline__ void oso_rotate( const double * sp, const double * sd, double * out )
{
    const double theta = sqrt( dot3( sd, sd ) );
    double sn, q;
    sincos( theta, &sn, &q );
    const double w = 1 - q, x = -sd[0] / theta, y = -sd[1] / theta, z = -sd[2] / theta;
    const double s1 = -y * z * w, s2 = x * z * w, s3 = -x * y * w, p1 = x * sn, p2 = y * sn, p3 = z * sn;
    if( theta > 1.0e-20 )
    {
        const double t1 = ( q + z * z * w ) * sp[0] + ( s1 + p1 ) * sp[1] + ( s2 + p2 ) * sp[2];
        const double t2 = ( s1 - p1 ) * sp[0] + ( q + y * y * w ) * sp[1] + ( s3 + p3 ) * sp[2];
        const double t3 = ( s2 - p2 ) * sp[0] + ( s3 - p3 ) * sp[1] + ( q + x * x * w ) * sp[2];
        out[0] = t1, out[1] = t2, out[2] = t3;
    }
    else
        out[0] = sp[0], out[1] = sp[1], out[2] = sp[2];
}

// Solver_Kernels::atlas_rotate (one spin)
__device__ __forceinline__ void atlas_rotate( const double * sp, double a3, const double * d, double * out )
{
    const double gamma = ( 1 + sp[2] * a3 );
    const double denom = ( sp[0] * sp[0] + sp[1] * sp[1] ) / gamma + 2 * ( d[0] * sp[0] + d[1] * sp[1] )
                         + gamma * ( d[0] * d[0] + d[1] * d[1] );
    double o0      = 2 * ( sp[0] + d[0] * gamma );
    double o1      = 2 * ( sp[1] + d[1] * gamma );
    double o2      = a3 * ( gamma - denom );
    const double r = 1 / ( gamma + denom );
    out[0] = o0 * r, out[1] = o1 * r, out[2] = o2 * r;
}

// ---------------------------------------------------------------------------
// force evaluation: energy partials, force, solver gradient, velocity (VP) or the newest
// L-BFGS pair and its dot products.  Chains take their forces from M.F_ext.
// ---------------------------------------------------------------------------
template<int SOLVER>
__global__ void __launch_bounds__( LLG_THREADS ) k_min_force( StepArgs A, MinArgs M, int nos )
{
    if( min_stopped( A, M ) )
        return;
    constexpr int D = SOLVER == SOLVER_LBFGS_ATLAS ? 2 : 3;
    const MinSite t = min_site( M.bpi, D, nos );
    const int i     = t.i;
    const int li    = M.sc->local_iter;
    double x[4]     = { 0, 0, 0, 0 }; // solver sums, then the energy
    if( i < nos )
    {
        double F[3];
        if( M.F_ext )
            load3( M.F_ext + t.b3, nos, i, F );
        else
        {
            const HamDev & H = A.imgs->ham;
            double g[3], e;
            local_gradient( H, A.s, nos, i, ( A.h && H.has_ddi ) ? A.h : nullptr, g, e );
            x[3] = e;
            F[0] = -1 * g[0], F[1] = -1 * g[1], F[2] = -1 * g[2];
            store3( M.f, nos, i, F );
        }
        double s[3];
        load3( A.s + t.b3, nos, i, s );
        if constexpr( SOLVER == SOLVER_VP || SOLVER == SOLVER_VP_OSO )
        {
            double gv[3];
            if constexpr( SOLVER == SOLVER_VP )
                gv[0] = F[0], gv[1] = F[1], gv[2] = F[2];
            else
            {
                double og[3];
                oso_gradient( s, F, og );
#pragma unroll
                for( int q = 0; q < 3; ++q )
                    gv[q] = og[q] * -1.0;
            }
            double pv[3], v[3];
            load3( M.prev + t.b3, nos, i, pv );
            load3( M.v + t.b3, nos, i, v );
#pragma unroll
            for( int q = 0; q < 3; ++q )
                v[q] += 0.5 / MIN_MASS * ( pv[q] + gv[q] );
            store3( M.v + t.b3, nos, i, v );
            store3( M.prev + t.b3, nos, i, gv );
            x[0] = dot3( v, gv );
            x[1] = dot3( gv, gv );
        }
        else
        {
            double gr[3];
            if constexpr( SOLVER == SOLVER_LBFGS_OSO )
                oso_gradient( s, F, gr );
            else
                atlas_gradient( s, F, M.a3[t.b1 + i], gr );
#pragma unroll
            for( int d = 0; d < D; ++d )
                M.grad[t.bd + d * nos + i] = gr[d];
            if( li > 0 )
            {
                const int m = li % MIN_MEMORY;
#pragma unroll
                for( int d = 0; d < D; ++d )
                {
                    const long long k = t.bd + d * nos + i;
                    const double a    = M.sd[k];
                    const double dgv  = gr[d] - M.grad_pr[k];
                    M.da[m][k]        = a;
                    M.dg[m][k]        = dgv;
                    x[0] += dgv * a;
                    x[1] += dgv * dgv;
                    x[2] += a * gr[d];
                }
            }
        }
    }
    min_block_sums<4>( x );
    if( threadIdx.x == 0 )
    {
        for( int s = 0; s < 3; ++s )
            M.part[s * M.n_blocks + blockIdx.x] = x[s];
        if( !M.F_ext )
            A.part_E[blockIdx.x] = x[3];
    }
}

// ---------------------------------------------------------------------------
// the scalar step between launches (one block)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__( SC_THREADS ) k_min_scalar( StepArgs A, MinArgs M, int phase, int round, int nos )
{
    if( min_stopped( A, M ) )
        return;
    __shared__ double sh[3][SC_THREADS];
    const int ns = ( phase == PH_RHO || phase == PH_FLIPRHO ) ? 3 : ( phase == PH_VP ? 2 : 1 );
    if( phase == PH_SCALE )
    {
        // rms of the search direction of each image (the whole block sums its partials),
        // then the largest
        double rmax = 0;
        for( int img = 0; img < M.noi; ++img )
        {
            double t = 0;
            for( int k = threadIdx.x; k < M.bpi; k += blockDim.x )
                t += M.part[(size_t)img * M.bpi + k];
            sh[0][threadIdx.x] = t;
            __syncthreads();
            for( int w = blockDim.x / 2; w > 0; w >>= 1 )
            {
                if( (int)threadIdx.x < w )
                    sh[0][threadIdx.x] += sh[0][threadIdx.x + w];
                __syncthreads();
            }
            rmax = fmax( rmax, sqrt( sh[0][0] / nos ) );
            __syncthreads();
        }
        if( threadIdx.x == 0 )
            sh[0][0] = rmax; // every thread holds the same value; thread 0 uses it below
    }
    else
    {
        for( int s = 0; s < ns; ++s )
        {
            double t = 0;
            for( int k = threadIdx.x; k < M.n_blocks; k += blockDim.x )
                t += M.part[s * M.n_blocks + k];
            sh[s][threadIdx.x] = t;
        }
        __syncthreads();
        for( int w = blockDim.x / 2; w > 0; w >>= 1 )
        {
            if( (int)threadIdx.x < w )
                for( int s = 0; s < ns; ++s )
                    sh[s][threadIdx.x] += sh[s][threadIdx.x + w];
            __syncthreads();
        }
    }
    if( threadIdx.x != 0 )
        return;
    MinScalars & S  = *M.sc;
    const double t0 = sh[0][0];
    const int li    = S.local_iter;
    switch( phase )
    {
        case PH_VP:
            S.ratio   = t0 / sh[1][0];
            S.vp_zero = t0 <= 0 ? 1 : 0;
            break;
        case PH_RHO:
        {
            const int m = li % MIN_MEMORY;
            int gd      = li == 0 ? 1 : 0;
            if( !gd )
            {
                if( t0 > LBFGS_EPS )
                {
                    S.rho[m]            = 1.0 / t0;
                    S.alpha[m]          = S.rho[m] * sh[2][0]; // first loop, k = M-1: c = m, q = grad
                    const double rhody2 = sh[1][0] * S.rho[m];
                    S.inv_rhody2        = rhody2 > LBFGS_EPS ? 1.0 / rhody2 : 1.0 / ( LBFGS_EPS );
                }
                else
                {
                    gd           = 1; // restart: steepest descent with an empty memory
                    S.local_iter = 0;
                }
            }
            if( gd )
                for( int k = 0; k < MIN_MEMORY; ++k )
                    S.rho[k] = 0.0;
            S.gd = gd;
            break;
        }
        case PH_ALPHA:
            if( !S.gd )
            {
                const int c = lbfgs_c1( li, round );
                S.alpha[c]  = S.rho[c] * t0;
            }
            break;
        case PH_COEF:
            if( !S.gd )
            {
                const int c   = lbfgs_c2( li, round );
                double rhopdg = t0;
                rhopdg *= S.rho[c];
                S.coef = S.alpha[c] - rhopdg;
            }
            break;
        case PH_SCALE:
            // t0: the largest rms over the images
            S.scaling    = t0 > M.maxmove ? M.maxmove / t0 : 1.0;
            S.local_iter = li + 1;
            break;
        case PH_FLIP: S.flip = t0 > 0 ? 1 : 0; break;
        case PH_FLIPRHO:
            if( S.flip )
                for( int n = 0; n < MIN_MEMORY; ++n )
                {
                    double r = 1 / S.rho[n];
                    r        = r + sh[n][0];
                    S.rho[n] = 1 / r;
                }
            break;
        default: break;
    }
}

// ---------------------------------------------------------------------------
// the two-loop recursion, one kernel per dependent dot product
// ---------------------------------------------------------------------------
// first loop, k = M-2 .. 0: q -= alpha_{c(k+1)} dg_{c(k+1)}, then da_{c(k)} . q
template<int D>
__global__ void __launch_bounds__( LLG_THREADS ) k_lbfgs_loop1( StepArgs A, MinArgs M, int nos, int k )
{
    if( min_stopped( A, M ) || M.sc->gd )
        return;
    const int li       = M.sc->local_iter;
    const int cp       = lbfgs_c1( li, k + 1 ), c = lbfgs_c1( li, k );
    const double ap    = M.sc->alpha[cp];
    const double * src = k == MIN_MEMORY - 2 ? M.grad : M.q;
    const MinSite t    = min_site( M.bpi, D, nos );
    double x[1]        = { 0 };
    if( t.i < nos )
    {
#pragma unroll
        for( int d = 0; d < D; ++d )
        {
            const long long j = t.bd + d * nos + t.i;
            double qv         = src[j];
            qv += -ap * M.dg[cp][j];
            M.q[j] = qv;
            x[0] += M.da[c][j] * qv;
        }
    }
    min_block_sums<1>( x );
    if( threadIdx.x == 0 )
        M.part[blockIdx.x] = x[0];
}

// end of the first loop and the scale: sd = (q - alpha_{c(0)} dg_{c(0)}) / (rho dy2); dg_{c'(0)} . sd
template<int D>
__global__ void __launch_bounds__( LLG_THREADS ) k_lbfgs_mid( StepArgs A, MinArgs M, int nos )
{
    if( min_stopped( A, M ) || M.sc->gd )
        return;
    const int li     = M.sc->local_iter;
    const int cp     = lbfgs_c1( li, 0 ), c = lbfgs_c2( li, 0 );
    const double ap  = M.sc->alpha[cp];
    const double inv = M.sc->inv_rhody2;
    const MinSite t  = min_site( M.bpi, D, nos );
    double x[1]      = { 0 };
    if( t.i < nos )
    {
#pragma unroll
        for( int d = 0; d < D; ++d )
        {
            const long long j = t.bd + d * nos + t.i;
            double qv         = M.q[j];
            qv += -ap * M.dg[cp][j];
            const double sdv = inv * qv;
            M.sd[j]          = sdv;
            x[0] += M.dg[c][j] * sdv;
        }
    }
    min_block_sums<1>( x );
    if( threadIdx.x == 0 )
        M.part[blockIdx.x] = x[0];
}

// second loop, k = 1 .. M-1: sd += coef_{k-1} da_{c'(k-1)}, then dg_{c'(k)} . sd
template<int D>
__global__ void __launch_bounds__( LLG_THREADS ) k_lbfgs_loop2( StepArgs A, MinArgs M, int nos, int k )
{
    if( min_stopped( A, M ) || M.sc->gd )
        return;
    const int li      = M.sc->local_iter;
    const int cp      = lbfgs_c2( li, k - 1 ), c = lbfgs_c2( li, k );
    const double coef = M.sc->coef;
    const MinSite t   = min_site( M.bpi, D, nos );
    double x[1]       = { 0 };
    if( t.i < nos )
    {
#pragma unroll
        for( int d = 0; d < D; ++d )
        {
            const long long j = t.bd + d * nos + t.i;
            double sdv        = M.sd[j];
            sdv += coef * M.da[cp][j];
            M.sd[j] = sdv;
            x[0] += M.dg[c][j] * sdv;
        }
    }
    min_block_sums<1>( x );
    if( threadIdx.x == 0 )
        M.part[blockIdx.x] = x[0];
}

// end of the recursion: sd = -(sd + coef da_{c'(M-1)}), grad_pr = grad; on the steepest-descent
// branch sd = -grad and the memory is cleared.  |sd|^2 partials for the step scaling.
template<int D>
__global__ void __launch_bounds__( LLG_THREADS ) k_lbfgs_final( StepArgs A, MinArgs M, int nos )
{
    if( min_stopped( A, M ) )
        return;
    const bool gd     = M.sc->gd;
    const int li      = M.sc->local_iter;
    const int cp      = lbfgs_c2( li, MIN_MEMORY - 1 );
    const double coef = M.sc->coef;
    const MinSite t   = min_site( M.bpi, D, nos );
    double x[1]       = { 0 };
    if( t.i < nos )
    {
        double sq = 0;
#pragma unroll
        for( int d = 0; d < D; ++d )
        {
            const long long j = t.bd + d * nos + t.i;
            const double g    = M.grad[j];
            double sdv;
            if( gd )
            {
                sdv = -g;
#pragma unroll
                for( int n = 0; n < MIN_MEMORY; ++n )
                {
                    M.da[n][j] = 0.0;
                    M.dg[n][j] = 0.0;
                }
            }
            else
            {
                sdv = M.sd[j];
                sdv += coef * M.da[cp][j];
                sdv = -sdv;
            }
            M.grad_pr[j] = g;
            M.sd[j]      = sdv;
            sq += sdv * sdv;
        }
        x[0] = sq;
    }
    min_block_sums<1>( x );
    if( threadIdx.x == 0 )
        M.part[blockIdx.x] = x[0];
}

// ---------------------------------------------------------------------------
// the update of the spins; for LLG runs Hook_Post_Iteration's torque / effective field
// ---------------------------------------------------------------------------
template<int SOLVER>
__global__ void __launch_bounds__( LLG_THREADS ) k_min_update( StepArgs A, MinArgs M, int nos )
{
    if( min_stopped( A, M ) )
        return;
    constexpr int D = SOLVER == SOLVER_LBFGS_ATLAS ? 2 : 3;
    const MinSite t = min_site( M.bpi, D, nos );
    const ImageDev * __restrict__ I   = A.imgs + t.img;
    const MinScalars * __restrict__ S = M.sc;
    const int i = t.i;
    double t_i  = 0;
    if( i < nos )
    {
        double s[3], sn[3];
        load3( A.s + t.b3, nos, i, s );
        const double dt = I->llg.dt;
        if constexpr( SOLVER == SOLVER_VP || SOLVER == SOLVER_VP_OSO )
        {
            double gv[3], v[3];
            load3( M.prev + t.b3, nos, i, gv );
#pragma unroll
            for( int q = 0; q < 3; ++q )
                v[q] = S->vp_zero ? 0.0 : gv[q] * S->ratio;
            store3( M.v + t.b3, nos, i, v );
            if constexpr( SOLVER == SOLVER_VP )
            {
#pragma unroll
                for( int q = 0; q < 3; ++q )
                    sn[q] = s[q] + dt * v[q] + 0.5 / MIN_MASS * dt * gv[q];
                normalize3( sn );
            }
            else
            {
                double sd[3];
#pragma unroll
                for( int q = 0; q < 3; ++q )
                    sd[q] = dt * v[q] + 0.5 / MIN_MASS * dt * gv[q];
                oso_rotate( s, sd, sn );
            }
        }
        else if constexpr( SOLVER == SOLVER_LBFGS_OSO )
        {
            double sd[3];
            load3( M.sd + t.bd, nos, i, sd );
#pragma unroll
            for( int q = 0; q < 3; ++q )
                sd[q] *= S->scaling;
            store3( M.sd + t.bd, nos, i, sd );
            oso_rotate( s, sd, sn );
        }
        else
        {
            double d[2] = { M.sd[t.bd + i], M.sd[t.bd + nos + i] };
            d[0] *= S->scaling;
            d[1] *= S->scaling;
            M.sd[t.bd + i]       = d[0];
            M.sd[t.bd + nos + i] = d[1];
            atlas_rotate( s, M.a3[t.b1 + i], d, sn );
        }
        store3( A.s + t.b3, nos, i, sn );
        if( !M.F_ext )
        {
            // Hook_Post_Iteration: torque of the virtual force (s x F at the old spins; scaled by
            // dt gamma / mu_B except for the L-BFGS solvers) projected on the new spins
            double F[3];
            load3( M.f, nos, i, F );
            constexpr bool LBFGS = SOLVER == SOLVER_LBFGS_OSO || SOLVER == SOLVER_LBFGS_ATLAS;
            const double c0      = LBFGS ? 1.0 : dt * gamma_e / mu_B;
            double c[3], fv[3], pr[3];
            cross3( s, F, c );
#pragma unroll
            for( int q = 0; q < 3; ++q )
                fv[q] = c0 * c[q];
            const double dd = dot3( fv, sn );
#pragma unroll
            for( int q = 0; q < 3; ++q )
                pr[q] = fv[q] - dd * sn[q];
            t_i = dot3( pr, pr );
            if( A.obs )
            {
                const double d1 = dot3( F, sn );
#pragma unroll
                for( int q = 0; q < 3; ++q )
                    pr[q] = F[q] - d1 * sn[q];
                store3( A.eff, nos, i, pr );
            }
        }
    }
    if( !M.F_ext )
    {
        const double tmax = min_block_max( t_i );
        if( threadIdx.x == 0 )
            A.part_T[blockIdx.x] = tmax;
    }
}

// ---------------------------------------------------------------------------
// atlas chart changes (ncg_atlas_check_coordinates, lbfgs_atlas_transform_direction)
// ---------------------------------------------------------------------------
// spins of image 0 against each image's charts, as the reference's check reads them
__global__ void __launch_bounds__( LLG_THREADS ) k_atlas_check( StepArgs A, MinArgs M, int nos )
{
    if( min_stopped( A, M ) )
        return;
    const MinSite t = min_site( M.bpi, 2, nos );
    double x[1]     = { 0 };
    if( t.i < nos )
        x[0] = A.s[2 * nos + t.i] * M.a3[t.b1 + t.i] < -0.6 ? 1.0 : 0.0;
    min_block_sums<1>( x );
    if( threadIdx.x == 0 )
        M.part[blockIdx.x] = x[0];
}

__device__ __forceinline__ bool atlas_leaves_chart( double sz, double a3, double & a3n, double & factor )
{
    if( !( sz * a3 < 0 ) )
        return false;
    a3n    = sz > 0 ? 1 : -1;
    factor = ( 1 - a3n * sz ) / ( 1 + a3n * sz );
    return true;
}

// per-spin terms (factor^2 - 1) da_n . dg_n of the spins that change chart (the reference
// adds them into rho_n from parallel threads; here they are summed in a fixed order)
__global__ void __launch_bounds__( LLG_THREADS ) k_atlas_flip_sums( StepArgs A, MinArgs M, int nos )
{
    if( min_stopped( A, M ) || !M.sc->flip )
        return;
    const MinSite t = min_site( M.bpi, 2, nos );
    double x[3]     = { 0, 0, 0 };
    double a3n, factor;
    if( t.i < nos && atlas_leaves_chart( A.s[t.b3 + 2 * nos + t.i], M.a3[t.b1 + t.i], a3n, factor ) )
    {
        const long long j0 = t.bd + t.i, j1 = t.bd + nos + t.i;
#pragma unroll
        for( int n = 0; n < MIN_MEMORY; ++n )
            x[n] = ( factor * factor - 1 ) * ( M.da[n][j0] * M.dg[n][j0] + M.da[n][j1] * M.dg[n][j1] );
    }
    min_block_sums<3>( x );
    if( threadIdx.x == 0 )
        for( int n = 0; n < MIN_MEMORY; ++n )
            M.part[n * M.n_blocks + blockIdx.x] = x[n];
}

__global__ void __launch_bounds__( LLG_THREADS ) k_atlas_flip_apply( StepArgs A, MinArgs M, int nos )
{
    if( min_stopped( A, M ) || !M.sc->flip )
        return;
    const MinSite t = min_site( M.bpi, 2, nos );
    double a3n, factor;
    if( t.i < nos && atlas_leaves_chart( A.s[t.b3 + 2 * nos + t.i], M.a3[t.b1 + t.i], a3n, factor ) )
    {
        M.a3[t.b1 + t.i] = a3n;
        for( int d = 0; d < 2; ++d )
        {
            const long long j = t.bd + d * nos + t.i;
            M.sd[j] *= factor;
            M.grad_pr[j] *= factor;
#pragma unroll
            for( int n = 0; n < MIN_MEMORY; ++n )
            {
                M.da[n][j] *= factor;
                M.dg[n][j] *= factor;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
bool min_is_solver( int solver )
{
    return solver == SOLVER_VP || solver == SOLVER_VP_OSO || solver == SOLVER_LBFGS_OSO
           || solver == SOLVER_LBFGS_ATLAS;
}

static void launch_scalar( const StepArgs & A, const MinArgs & M, int phase, int round, int nos, hipStream_t st )
{
    hipLaunchKernelGGL( k_min_scalar, dim3( 1 ), dim3( SC_THREADS ), 0, st, A, M, phase, round, nos );
}

template<int SOLVER>
static void vp_iteration( const StepArgs & A, const MinArgs & M, int nos, hipStream_t st )
{
    const dim3 g( M.n_blocks ), b( LLG_THREADS );
    hipLaunchKernelGGL( k_min_force<SOLVER>, g, b, 0, st, A, M, nos );
    launch_scalar( A, M, PH_VP, 0, nos, st );
    hipLaunchKernelGGL( k_min_update<SOLVER>, g, b, 0, st, A, M, nos );
}

template<int SOLVER>
static void lbfgs_iteration( const StepArgs & A, const MinArgs & M, int nos, hipStream_t st )
{
    constexpr int D = SOLVER == SOLVER_LBFGS_ATLAS ? 2 : 3;
    const dim3 g( M.n_blocks ), b( LLG_THREADS );
    hipLaunchKernelGGL( k_min_force<SOLVER>, g, b, 0, st, A, M, nos );
    launch_scalar( A, M, PH_RHO, 0, nos, st );
    for( int k = MIN_MEMORY - 2; k >= 0; --k )
    {
        hipLaunchKernelGGL( k_lbfgs_loop1<D>, g, b, 0, st, A, M, nos, k );
        launch_scalar( A, M, PH_ALPHA, k, nos, st );
    }
    hipLaunchKernelGGL( k_lbfgs_mid<D>, g, b, 0, st, A, M, nos );
    launch_scalar( A, M, PH_COEF, 0, nos, st );
    for( int k = 1; k < MIN_MEMORY; ++k )
    {
        hipLaunchKernelGGL( k_lbfgs_loop2<D>, g, b, 0, st, A, M, nos, k );
        launch_scalar( A, M, PH_COEF, k, nos, st );
    }
    hipLaunchKernelGGL( k_lbfgs_final<D>, g, b, 0, st, A, M, nos );
    launch_scalar( A, M, PH_SCALE, 0, nos, st );
    hipLaunchKernelGGL( k_min_update<SOLVER>, g, b, 0, st, A, M, nos );
    if constexpr( SOLVER == SOLVER_LBFGS_ATLAS )
    {
        hipLaunchKernelGGL( k_atlas_check, g, b, 0, st, A, M, nos );
        launch_scalar( A, M, PH_FLIP, 0, nos, st );
        hipLaunchKernelGGL( k_atlas_flip_sums, g, b, 0, st, A, M, nos );
        launch_scalar( A, M, PH_FLIPRHO, 0, nos, st );
        hipLaunchKernelGGL( k_atlas_flip_apply, g, b, 0, st, A, M, nos );
    }
}

void min_launch_chain_update( int solver, const StepArgs & A, const MinArgs & M, int nos, hipStream_t st )
{
    switch( solver )
    {
        case SOLVER_VP_OSO: vp_iteration<SOLVER_VP_OSO>( A, M, nos, st ); break;
        case SOLVER_LBFGS_OSO: lbfgs_iteration<SOLVER_LBFGS_OSO>( A, M, nos, st ); break;
        case SOLVER_LBFGS_ATLAS: lbfgs_iteration<SOLVER_LBFGS_ATLAS>( A, M, nos, st ); break;
        default: break;
    }
}

void min_launch_iteration( int solver, const StepArgs & A, const MinArgs & M, int nos, hipStream_t st )
{
    switch( solver )
    {
        case SOLVER_VP: vp_iteration<SOLVER_VP>( A, M, nos, st ); break;
        case SOLVER_VP_OSO: vp_iteration<SOLVER_VP_OSO>( A, M, nos, st ); break;
        case SOLVER_LBFGS_OSO: lbfgs_iteration<SOLVER_LBFGS_OSO>( A, M, nos, st ); break;
        case SOLVER_LBFGS_ATLAS: lbfgs_iteration<SOLVER_LBFGS_ATLAS>( A, M, nos, st ); break;
        default: break;
    }
}

} // namespace spirit_hip
