// Metropolis Monte Carlo (FP64), hand-written for gfx950: colour classes of spins that share
// no pair are updated in parallel; each trial move, energy difference and acceptance test is
// one thread's work, with counter-based random numbers (Philox4x32-10) keyed by the seed,
// step, colour and spin.
//
// Restated reference code (paths relative to the reference checkout):
//   Method_MC::Metropolis (cone sampling, local basis, acceptance, cone feedback)
//                                             core/src/engine/Method_MC.cpp:62-185
//   Hamiltonian_Heisenberg::Energy_Single_Spin core/src/engine/Hamiltonian_Heisenberg.cpp:559-668
#include "llg_device.hpp"
#include "mc.hpp"

namespace spirit_hip
{

constexpr int MC_THREADS = 256;
constexpr double MC_CONE_STEP = 0.01; // Method_MC.cpp:70 (diff)

// Energy_Single_Spin of spin i (basis atom ib in cell a, b, c) oriented along s, the other
// spins as in x (redundant pair tables: each pair of the spin once, as the OpenMP build)
// G: se holds the exchange neighbours gathered into registers (gather_exchange_neighbours,
// zeros where a pair is cut); DMI uses them when its table coincides with exchange's
template<bool G>
__device__ __forceinline__ double single_spin_energy( const HamDev & H, const double * __restrict__ x, int nos, int i,
                                                      int ib, int a, int b, int c, const double * s,
                                                      const double ( &se )[LG_MAX_NB][3] )
{
    double E = 0;
    if( H.has_zeeman )
        E -= H.mu[ib] * H.B * dot3( H.Bn, s );
    for( int k = 0; k < H.n_aniso; ++k )
        if( H.aniso_idx[k] == ib )
            E -= H.aniso_K[k] * pow( dot3( H.aniso_n[k], s ), 2.0 );
    if( H.has_site_K )
        E -= H.site_K[i] * pow( dot3( H.site_n, s ), 2.0 );
    for( int k = 0; k < H.n_cubic; ++k )
        if( H.cubic_idx[k] == ib )
            E -= H.cubic_K[k] / 2 * ( pow( s[0], 4.0 ) + pow( s[1], 4.0 ) + pow( s[2], 4.0 ) );
    if constexpr( G )
    {
#pragma unroll
        for( int k = 0; k < LG_MAX_NB; ++k )
            if( k < H.n_exch )
                E -= H.exch[k].J * dot3( s, se[k] );
    }
    else
        for( int k = 0; k < H.n_exch; ++k )
        {
            const PairDev & p = H.exch[k];
            if( p.i != ib )
                continue;
            const int j = pair_neighbour( H, a, b, c, p );
            if( j >= 0 )
            {
                const double sj[3] = { x[j], x[nos + j], x[2 * nos + j] };
                E -= p.J * dot3( s, sj );
            }
        }
    if( G && H.dmi_same )
    {
#pragma unroll
        for( int k = 0; k < LG_MAX_NB; ++k )
            if( k < H.n_dmi )
            {
                const PairDev & p = H.dmi[k];
                double cr[3];
                cross3( s, se[k], cr );
                E -= p.D * dot3( p.n, cr );
            }
        return E;
    }
    for( int k = 0; k < H.n_dmi; ++k )
    {
        const PairDev & p = H.dmi[k];
        if( p.i != ib )
            continue;
        const int j = pair_neighbour( H, a, b, c, p );
        if( j >= 0 )
        {
            const double sj[3] = { x[j], x[nos + j], x[2 * nos + j] };
            double cr[3];
            cross3( s, sj, cr );
            E -= p.D * dot3( p.n, cr );
        }
    }
    return E;
}

// four uniform doubles in [0, 1) for (seed, step, colour, spin)
__device__ __forceinline__ void mc_uniforms( uint64_t seed, uint64_t step, int colour, int i, double * u )
{
    uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)( seed >> 32 );
    uint32_t c0[4] = { (uint32_t)i, (uint32_t)colour | 0x40000000u, (uint32_t)step, (uint32_t)( step >> 32 ) };
    uint32_t c1[4] = { c0[0], c0[1] | 0x80000000u, c0[2], c0[3] };
    philox( c0, k0, k1 );
    philox( c1, k0, k1 );
    u[0] = u53( c0[0], c0[1] );
    u[1] = u53( c0[2], c0[3] );
    u[2] = u53( c1[0], c1[1] );
    u[3] = u53( c1[2], c1[3] );
}

// Metropolis trial of site i (cell (a, b, c), basis atom ib) of colour class `colour`;
// returns 1 when the move is rejected.  McArgs is taken by value so that its fields stay
// kernel arguments (scalar loads).
__device__ __forceinline__ double mc_site( McArgs A, int colour, int i, int ib, int a, int b, int c )
{
    const HamDev & H = A.img->ham;
    const int nos    = A.nos;
    // the neighbours' loads go out first: they do not depend on the trial move
    const bool gather = H.n_exch <= LG_MAX_NB;
    double se[LG_MAX_NB][3];
    if( gather )
        gather_exchange_neighbours( H, A.s, nos, i, ib, a, b, c, se );
    double u[4];
    mc_uniforms( A.seed, *A.step, colour, i, u );
    const double s[3] = { A.s[i], A.s[nos + i], A.s[2 * nos + i] };
    double sn[3];
    if( A.step_cone )
    {
        // local basis (e0, e1, s) of the spin, Method_MC.cpp:116-126
        double e0[3] = { 1, 0, 0 }, e1[3] = { 0, 1, 0 }, e2[3] = { 0, 0, 1 };
        if( s[2] < 1 - 1e-10 )
        {
            const double ez[3] = { 0, 0, 1 };
            e2[0] = s[0], e2[1] = s[1], e2[2] = s[2];
            cross3( e2, ez, e0 );
            normalize3( e0 );
            cross3( e2, e0, e1 );
        }
        const double cos_cone = A.cone[1];
        const double costheta = 1 - ( 1 - cos_cone ) * u[0];
        const double sintheta = sqrt( 1 - costheta * costheta );
        const double phi      = 2 * Pi * u[1];
        double sp, cp;
        sincos( phi, &sp, &cp );
        const double l[3] = { sintheta * cp, sintheta * sp, costheta };
#pragma unroll
        for( int q = 0; q < 3; ++q )
            sn[q] = e0[q] * l[0] + e1[q] * l[1] + e2[q] * l[2];
    }
    else
    {
        // Method_MC.cpp:140-149: cos(theta) drawn from [0, 1)
        const double costheta = u[0];
        const double sintheta = sqrt( 1 - costheta * costheta );
        const double phi      = 2 * Pi * u[1];
        double sp, cp;
        sincos( phi, &sp, &cp );
        sn[0] = sintheta * cp, sn[1] = sintheta * sp, sn[2] = costheta;
    }
    double Eold, Enew;
    if( gather )
    {
        Eold = single_spin_energy<true>( H, A.s, nos, i, ib, a, b, c, s, se );
        Enew = single_spin_energy<true>( H, A.s, nos, i, ib, a, b, c, sn, se );
    }
    else
    {
        Eold = single_spin_energy<false>( H, A.s, nos, i, ib, a, b, c, s, se );
        Enew = single_spin_energy<false>( H, A.s, nos, i, ib, a, b, c, sn, se );
    }
    const double Ediff = Enew - Eold;
    bool accept        = true;
    if( Ediff > 1e-14 )
    {
        if( A.temperature < 1e-12 )
            accept = false;
        else if( exp( -Ediff / A.kT ) < u[2] )
            accept = false;
    }
    if( accept )
    {
        A.s[i] = sn[0], A.s[nos + i] = sn[1], A.s[2 * nos + i] = sn[2];
    }
    else
        return 1;
    return 0;
}

// block-wide rejection count in a fixed order into part[colour][blockIdx.x]
__device__ __forceinline__ void mc_block_rejections( const McArgs & A, int colour, double rejected )
{
    __shared__ double red[MC_THREADS / 64];
#pragma unroll
    for( int off = 32; off > 0; off >>= 1 )
        rejected += __shfl_down( rejected, off, 64 );
    if( threadIdx.x % 64 == 0 )
        red[threadIdx.x / 64] = rejected;
    __syncthreads();
    if( threadIdx.x == 0 )
    {
        double t = red[0];
        for( int k = 1; k < MC_THREADS / 64; ++k )
            t += red[k];
        A.part[colour * A.n_blocks + blockIdx.x] = t;
    }
}

// one thread per site; the sites of other colours idle
__global__ void __launch_bounds__( MC_THREADS ) k_mc_colour( McArgs A, int colour )
{
    const HamDev & H = A.img->ham;
    const int i      = blockIdx.x * MC_THREADS + threadIdx.x;
    double rejected  = 0;
    if( i < A.nos )
    {
        const int cell = fast_div( i, H.div_nb );
        const int ib   = i - cell * H.nb;
        const int rest = fast_div( cell, H.div_Na );
        const int a    = cell - rest * H.Na;
        const int c    = fast_div( rest, H.div_Nb );
        const int b    = rest - c * H.Nb;
        const int col  = ( A.ca * a + A.cb * b + A.cc * c ) % A.K;
        if( col == colour )
            rejected = mc_site( A, colour, i, ib, a, b, c );
    }
    mc_block_rejections( A, colour, rejected );
}

// one thread per site of the colour (one basis atom, ca invertible mod K): along a row
// (b, c) the colour repeats with period K, so group g of K consecutive cells holds exactly
// one site of each colour, at a = K g + inv_ca (colour - cb b - cc c) mod K
__global__ void __launch_bounds__( MC_THREADS ) k_mc_colour_rows( McArgs A, int colour )
{
    const HamDev & H = A.img->ham;
    const int t      = blockIdx.x * MC_THREADS + threadIdx.x;
    const int row    = t / A.groups;
    double rejected  = 0;
    if( row < H.Nb * H.Nc )
    {
        const int g  = t - row * A.groups;
        const int c  = row / H.Nb;
        const int b  = row - c * H.Nb;
        const int K  = A.K;
        const int a0 = ( ( A.inv_ca * ( colour - ( A.cb * b + A.cc * c ) % K ) ) % K + K ) % K;
        const int a  = K * g + a0;
        if( a < H.Na )
            rejected = mc_site( A, colour, a + H.Na * row, 0, a, b, c );
    }
    mc_block_rejections( A, colour, rejected );
}

// Cone angle feedback from the rejections of the previous iteration (Method_MC.cpp:66-79)
__global__ void __launch_bounds__( 1024 ) k_mc_cone( McArgs A )
{
    __shared__ double sh[1024];
    double t = 0;
    for( int k = threadIdx.x; k < A.K * A.n_blocks; k += blockDim.x )
        t += A.part[k];
    sh[threadIdx.x] = t;
    __syncthreads();
    for( int w = blockDim.x / 2; w > 0; w >>= 1 )
    {
        if( (int)threadIdx.x < w )
            sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if( threadIdx.x != 0 )
        return;
    double cone = A.cone[0];
    if( A.step_cone && A.adaptive )
    {
        const double acc = 1 - sh[0] / (double)A.nos;
        if( acc < A.target && cone > MC_CONE_STEP )
            cone -= MC_CONE_STEP;
        if( acc > A.target && cone < Pi - MC_CONE_STEP )
            cone += MC_CONE_STEP;
        A.cone[0] = cone;
    }
    A.cone[1] = cos( cone ); // for every trial move of the iteration
}

__global__ void k_mc_advance( uint64_t * step )
{
    if( threadIdx.x == 0 )
        *step += 1;
}

void mc_launch_iteration( const McArgs & A, uint64_t * step, hipStream_t st )
{
    hipLaunchKernelGGL( k_mc_cone, dim3( 1 ), dim3( 1024 ), 0, st, A );
    for( int c = 0; c < A.K; ++c )
    {
        if( A.inv_ca > 0 )
            hipLaunchKernelGGL( k_mc_colour_rows, dim3( A.n_blocks_rows ), dim3( MC_THREADS ), 0, st, A, c );
        else
            hipLaunchKernelGGL( k_mc_colour, dim3( A.n_blocks ), dim3( MC_THREADS ), 0, st, A, c );
    }
    hipLaunchKernelGGL( k_mc_advance, dim3( 1 ), dim3( 64 ), 0, st, step );
}

} // namespace spirit_hip
