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
#include "llg.hpp"

namespace spirit_hip
{

// ---------------------------------------------------------------------------
// small vector helpers with the reference's (Eigen's) evaluation order
// ---------------------------------------------------------------------------
__device__ __forceinline__ double dot3( const double * a, const double * b )
{
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}
__device__ __forceinline__ void cross3( const double * a, const double * b, double * o )
{
    double x = a[1] * b[2] - a[2] * b[1];
    double y = a[2] * b[0] - a[0] * b[2];
    double z = a[0] * b[1] - a[1] * b[0];
    o[0] = x, o[1] = y, o[2] = z;
}
__device__ __forceinline__ void normalize3( double * a )
{
    double z = dot3( a, a );
    if( z > 0 )
    {
        double n = sqrt( z );
        a[0] /= n;
        a[1] /= n;
        a[2] /= n;
    }
}
// Rodrigues rotation, Vectormath.cpp:474-479
__device__ __forceinline__ void rotate3( const double * v, const double * axis, double angle, double * out )
{
    double sn, c;
    sincos( angle, &sn, &c );
    double cr[3];
    cross3( axis, v, cr );
    double d = dot3( axis, v );
    double o[3];
#pragma unroll
    for( int q = 0; q < 3; ++q )
        o[q] = v[q] * c + cr[q] * sn + axis[q] * d * ( 1 - c );
    out[0] = o[0], out[1] = o[1], out[2] = o[2];
}
// sib_transform, Solver_Kernels.cpp:15-48
__device__ __forceinline__ void sib3( const double * s, const double * f, double * out )
{
    double A[3] = { 0.5 * f[0], 0.5 * f[1], 0.5 * f[2] };
    double an   = sqrt( dot3( A, A ) );
    double detAi = 1.0 / ( 1 + an * an );
    double c[3], a2[3];
    cross3( s, A, c );
#pragma unroll
    for( int q = 0; q < 3; ++q )
        a2[q] = s[q] - c[q];
    double o0 = ( a2[0] * ( A[0] * A[0] + 1 ) + a2[1] * ( A[0] * A[1] - A[2] ) + a2[2] * ( A[0] * A[2] + A[1] ) ) * detAi;
    double o1 = ( a2[0] * ( A[1] * A[0] + A[2] ) + a2[1] * ( A[1] * A[1] + 1 ) + a2[2] * ( A[1] * A[2] - A[0] ) ) * detAi;
    double o2 = ( a2[0] * ( A[2] * A[0] - A[1] ) + a2[1] * ( A[2] * A[1] + A[0] ) + a2[2] * ( A[2] * A[2] + 1 ) ) * detAi;
    out[0] = o0, out[1] = o1, out[2] = o2;
}
__device__ __forceinline__ void neg_cross3( const double * a, const double * b, double * o )
{
    double c[3];
    cross3( a, b, c );
    o[0] = -1 * c[0], o[1] = -1 * c[1], o[2] = -1 * c[2];
}

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon et al., SC'11) + Box-Muller: three N(0,1) per (step, site)
// ---------------------------------------------------------------------------
__device__ __forceinline__ void philox( uint32_t c[4], uint32_t k0, uint32_t k1 )
{
#pragma unroll
    for( int r = 0; r < 10; ++r )
    {
        uint32_t hi0 = __umulhi( 0xD2511F53u, c[0] ), lo0 = 0xD2511F53u * c[0];
        uint32_t hi1 = __umulhi( 0xCD9E8D57u, c[2] ), lo1 = 0xCD9E8D57u * c[2];
        uint32_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
        c[0] = n0, c[1] = lo1, c[2] = n2, c[3] = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}
__device__ __forceinline__ double u53( uint32_t a, uint32_t b )
{
    return ( (double)( a >> 5 ) * 67108864.0 + (double)( b >> 6 ) ) * ( 1.0 / 9007199254740992.0 );
}
__device__ __forceinline__ void normal3( uint64_t seed, uint64_t step, uint64_t site, double * z )
{
    uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)( seed >> 32 );
    uint32_t c0[4] = { (uint32_t)site, (uint32_t)( site >> 32 ), (uint32_t)step, (uint32_t)( step >> 32 ) };
    uint32_t c1[4] = { c0[0], c0[1] | 0x80000000u, c0[2], c0[3] };
    philox( c0, k0, k1 );
    philox( c1, k0, k1 );
    double u1 = 1.0 - u53( c0[0], c0[1] );
    double u2 = u53( c0[2], c0[3] );
    double u3 = 1.0 - u53( c1[0], c1[1] );
    double u4 = u53( c1[2], c1[3] );
    double r1 = sqrt( -2.0 * log( u1 ) );
    double r2 = sqrt( -2.0 * log( u3 ) );
    z[0]      = r1 * cos( 2.0 * Pi * u2 );
    z[1]      = r1 * sin( 2.0 * Pi * u2 );
    z[2]      = r2 * cos( 2.0 * Pi * u4 );
}

// ---------------------------------------------------------------------------
// neighbour index, Vectormath.hpp:437-528 (defects off)
// ---------------------------------------------------------------------------
__device__ __forceinline__ int pair_neighbour( const HamDev & H, int a, int b, int c, const PairDev & p )
{
    const int Na = H.Na, Nb = H.Nb, Nc = H.Nc;
    if( abs( p.da ) > Na || abs( p.db ) > Nb || abs( p.dc ) > Nc )
        return -1;
    int ja = a + p.da, jb = b + p.db, jc = c + p.dc;
    if( H.bc[0] || ( 0 <= ja && ja < Na ) )
    {
        if( ja < 0 )
            ja += Na;
        if( ja >= Na )
            ja -= Na;
    }
    else
        return -1;
    if( H.bc[1] || ( 0 <= jb && jb < Nb ) )
    {
        if( jb < 0 )
            jb += Nb;
        if( jb >= Nb )
            jb -= Nb;
    }
    else
        return -1;
    if( H.bc[2] || ( 0 <= jc && jc < Nc ) )
    {
        if( jc < 0 )
            jc += Nc;
        if( jc >= Nc )
            jc -= Nc;
    }
    else
        return -1;
    return p.j + H.nb * ( ja + Na * ( jb + Nb * jc ) );
}

// Gradient_and_Energy for one spin: g (all terms) and its energy share e
// (1/2 g.s over aniso+exchange+DMI+DDI, plus the cubic and Zeeman energies).
__device__ __forceinline__ void local_gradient( const HamDev & H, const double * __restrict__ x, int nos, int i,
                                                const double * __restrict__ hddi, double * g, double & e )
{
    const int ib   = i % H.nb;
    const int cell = i / H.nb;
    const int a    = cell % H.Na;
    const int rest = cell / H.Na;
    const int b    = rest % H.Nb;
    const int c    = rest / H.Nb;
    const double s[3] = { x[i], x[nos + i], x[2 * nos + i] };
    g[0] = g[1] = g[2] = 0;
    // anisotropy, :785-800
    for( int k = 0; k < H.n_aniso; ++k )
    {
        if( H.aniso_idx[k] != ib )
            continue;
        const double * n = H.aniso_n[k];
        double d         = dot3( n, s );
#pragma unroll
        for( int q = 0; q < 3; ++q )
            g[q] -= 2.0 * H.aniso_K[k] * n[q] * d;
    }
    if( H.has_site_K )
    {
        double d = dot3( H.site_n, s );
#pragma unroll
        for( int q = 0; q < 3; ++q )
            g[q] -= 2.0 * H.site_K[i] * H.site_n[q] * d;
    }
    // exchange, :822-842 (gather over the redundant table)
    for( int k = 0; k < H.n_exch; ++k )
    {
        const PairDev & p = H.exch[k];
        if( p.i != ib )
            continue;
        int j = pair_neighbour( H, a, b, c, p );
        if( j >= 0 )
        {
            g[0] -= p.J * x[j];
            g[1] -= p.J * x[nos + j];
            g[2] -= p.J * x[2 * nos + j];
        }
    }
    // DMI, :844-864
    for( int k = 0; k < H.n_dmi; ++k )
    {
        const PairDev & p = H.dmi[k];
        if( p.i != ib )
            continue;
        int j = pair_neighbour( H, a, b, c, p );
        if( j >= 0 )
        {
            const double sj[3] = { x[j], x[nos + j], x[2 * nos + j] };
            double cr[3];
            cross3( sj, p.n, cr );
#pragma unroll
            for( int q = 0; q < 3; ++q )
                g[q] -= p.D * cr[q];
        }
    }
    // DDI field (precomputed by the FFT passes), :995-1013
    if( hddi )
    {
        g[0] -= hddi[i];
        g[1] -= hddi[nos + i];
        g[2] -= hddi[2 * nos + i];
    }
    e = 0.5 * dot3( g, s );
    // cubic anisotropy, :802-820 and :342-358
    for( int k = 0; k < H.n_cubic; ++k )
    {
        if( H.cubic_idx[k] != ib )
            continue;
#pragma unroll
        for( int q = 0; q < 3; ++q )
            g[q] -= 2.0 * H.cubic_K[k] * pow( s[q], 3.0 );
        e -= H.cubic_K[k] / 2 * ( pow( s[0], 4.0 ) + pow( s[1], 4.0 ) + pow( s[2], 4.0 ) );
    }
    // Zeeman, :768-783 and :745-751
    if( H.has_zeeman )
    {
        const double mu = H.mu[ib];
#pragma unroll
        for( int q = 0; q < 3; ++q )
            g[q] -= mu * H.B * H.Bn[q];
        double ext[3] = { H.Bn[0] * H.B, H.Bn[1] * H.B, H.Bn[2] * H.B };
        e += -mu * dot3( ext, s );
    }
}

// Method_LLG::Calculate_Force_Virtual (no STT), Method_LLG.cpp:132-226
__device__ __forceinline__ void virtual_force( const LLGDev & P, int solver, double mu, const double * s,
                                               const double * F, const double * xi, double * fv )
{
    const double a = P.damping;
    if( P.direct_minimization || solver == SOLVER_VP )
    {
        const double dtg = P.dt * gamma_e / mu_B;
        double c[3];
        cross3( s, F, c );
#pragma unroll
        for( int q = 0; q < 3; ++q )
            fv[q] = dtg * c[q];
        return;
    }
    const double dtg = P.dt * gamma_e / mu_B / ( 1 + a * a );
    double c[3], v[3];
#pragma unroll
    for( int q = 0; q < 3; ++q )
        v[q] = dtg * F[q];
    cross3( s, F, c );
#pragma unroll
    for( int q = 0; q < 3; ++q )
        v[q] += dtg * a * c[q];
#pragma unroll
    for( int q = 0; q < 3; ++q )
        v[q] /= mu;
    if( P.thermal )
    {
#pragma unroll
        for( int q = 0; q < 3; ++q )
            v[q] += 1 * xi[q];
        cross3( s, xi, c );
#pragma unroll
        for( int q = 0; q < 3; ++q )
            v[q] += a * c[q];
    }
    fv[0] = v[0], fv[1] = v[1], fv[2] = v[2];
}

// thermal field xi_i = eps * sqrt(T/mu_i) * N(0,1), Method_LLG.cpp:66-110
__device__ __forceinline__ void thermal_field( const LLGDev & P, double mu, uint64_t step, int i,
                                               const double * __restrict__ noise, double * xi )
{
    double z[3];
    if( noise )
    {
        z[0] = noise[3LL * i + 0];
        z[1] = noise[3LL * i + 1];
        z[2] = noise[3LL * i + 2];
    }
    else
        normal3( P.seed, step, (uint64_t)i, z );
    const double a   = P.damping;
    const double eps = sqrt( 2 * a * P.dt * gamma_e / mu_B * k_B ) / ( 1 + a * a );
    const double f   = eps * sqrt( P.temperature / mu );
#pragma unroll
    for( int q = 0; q < 3; ++q )
        xi[q] = f * z[q];
}

// block-wide reduction of (sum e, max t) into partials[blockIdx.x]
__device__ __forceinline__ void block_reduce_store( double e, double t, double * part_E, double * part_T )
{
    __shared__ double sE[LLG_THREADS / 64], sT[LLG_THREADS / 64];
#pragma unroll
    for( int off = 32; off > 0; off >>= 1 )
    {
        e += __shfl_down( e, off, 64 );
        t = fmax( t, __shfl_down( t, off, 64 ) );
    }
    const int w = threadIdx.x / 64, lane = threadIdx.x % 64;
    if( lane == 0 )
    {
        sE[w] = e;
        sT[w] = t;
    }
    __syncthreads();
    if( threadIdx.x == 0 )
    {
        double E = sE[0], T = sT[0];
        for( int k = 1; k < LLG_THREADS / 64; ++k )
        {
            E += sE[k];
            T = fmax( T, sT[k] );
        }
        part_E[blockIdx.x] = E;
        part_T[blockIdx.x] = T;
    }
}

__device__ __forceinline__ void load3( const double * f, int nos, int i, double * v )
{
    v[0] = f[i], v[1] = f[nos + i], v[2] = f[2 * nos + i];
}
__device__ __forceinline__ void store3( double * f, int nos, int i, const double * v )
{
    f[i] = v[0], f[nos + i] = v[1], f[2 * nos + i] = v[2];
}

// ---------------------------------------------------------------------------
// the stage kernel
// ---------------------------------------------------------------------------
template<int SOLVER, int STAGE>
__global__ void __launch_bounds__( LLG_THREADS ) k_llg_stage( StepArgs A )
{
    if( *A.stop )
        return;
    const int blk           = blockIdx.x;
    const ImageDev * __restrict__ I = A.imgs + A.blk_img[blk];
    const HamDev & H        = I->ham;
    const int nos           = H.nos;
    const long long base    = 3 * I->off;
    const int i             = ( blk - I->block_begin ) * LLG_THREADS + threadIdx.x;
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
        double g[3], e;
        local_gradient( H, x, nos, i, ( A.h && H.has_ddi ) ? A.h + base : nullptr, g, e );
        e_i            = e;
        const double mu = H.mu[i % H.nb];
        double si[3];
        load3( x, nos, i, si );
        const double F[3] = { -1 * g[0], -1 * g[1], -1 * g[2] };
        double xi[3]      = { 0, 0, 0 };
        if( I->llg.thermal && !I->llg.direct_minimization )
            thermal_field( I->llg, mu, A.step, i, A.noise ? A.noise + base : nullptr, xi );
        double fv[3];
        virtual_force( I->llg, SOLVER, mu, si, F, xi, fv );
        if( STAGE == 0 && A.obs )
            store3( A.eff_tmp + base, nos, i, F );

        double s0[3];
        if( STAGE == 0 )
        {
            s0[0] = si[0], s0[1] = si[1], s0[2] = si[2];
        }
        else
            load3( A.s + base, nos, i, s0 );
        double snew[3];
        double fv1[3];

        if( SOLVER == SOLVER_DEPONDT )
        {
            if( STAGE == 0 )
            {
                store3( A.aux + base, nos, i, fv );
                double axis[3] = { 1 * fv[0], 1 * fv[1], 1 * fv[2] };
                double angle   = sqrt( dot3( fv, fv ) );
                normalize3( axis );
                double pr[3];
                rotate3( si, axis, angle, pr );
                store3( A.sp + base, nos, i, pr );
            }
            else
            {
                load3( A.aux + base, nos, i, fv1 );
                double t[3];
#pragma unroll
                for( int q = 0; q < 3; ++q )
                    t[q] = 0.5 * fv1[q];
#pragma unroll
                for( int q = 0; q < 3; ++q )
                    t[q] += 0.5 * fv[q];
                double angle = sqrt( dot3( t, t ) );
                normalize3( t );
                rotate3( s0, t, angle, snew );
                store3( A.s + base, nos, i, snew );
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
                normalize3( pr );
                store3( A.sp + base, nos, i, pr );
            }
            else
            {
                load3( A.aux + base, nos, i, fv1 );
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
                normalize3( t );
                snew[0] = t[0], snew[1] = t[1], snew[2] = t[2];
                store3( A.s + base, nos, i, snew );
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
                store3( A.s + base, nos, i, snew );
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
                normalize3( pr );
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
                normalize3( pr );
                store3( ( STAGE == 1 ? A.sp2 : A.sp ) + base, nos, i, pr );
            }
            else
            {
                double acc[3];
                load3( A.aux + base, nos, i, acc );
#pragma unroll
                for( int q = 0; q < 3; ++q )
                    acc[q] += 1.0 / 6.0 * k[q];
                normalize3( acc );
                snew[0] = acc[0], snew[1] = acc[1], snew[2] = acc[2];
                store3( A.s + base, nos, i, snew );
                if( A.obs )
                    load3( A.aux2 + base, nos, i, fv1 );
            }
        }

        if( LAST && A.obs )
        {
            // MaxTorque_on_Image: |P_perp(s_new) F_v1|; effective field P_perp(s_new) F1
            double d = dot3( fv1, snew );
            double pr[3];
#pragma unroll
            for( int q = 0; q < 3; ++q )
                pr[q] = fv1[q] - d * snew[q];
            t_i = dot3( pr, pr );
            double F1[3];
            load3( A.eff_tmp + base, nos, i, F1 );
            double d1 = dot3( F1, snew );
#pragma unroll
            for( int q = 0; q < 3; ++q )
                pr[q] = F1[q] - d1 * snew[q];
            store3( A.eff + base, nos, i, pr );
        }
    }
    if( LAST && A.obs )
        block_reduce_store( e_i, t_i, A.part_E, A.part_T );
}

// The method constructor's force evaluation and Hook_Post_Iteration (Method_LLG.cpp:22-63,
// 247-301) on the current spins, without a thermal field: energy, max torque of the
// virtual force projected on s, effective field P_perp(s) F.
__global__ void __launch_bounds__( LLG_THREADS ) k_observe( StepArgs A, int solver )
{
    const int blk           = blockIdx.x;
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
        double fv[3];
        virtual_force( P, solver, H.mu[i % H.nb], si, F, xi, fv );
        double d = dot3( fv, si ), pr[3];
#pragma unroll
        for( int q = 0; q < 3; ++q )
            pr[q] = fv[q] - d * si[q];
        t_i       = dot3( pr, pr );
        double d1 = dot3( F, si );
#pragma unroll
        for( int q = 0; q < 3; ++q )
            pr[q] = F[q] - d1 * si[q];
        store3( A.eff + base, nos, i, pr );
    }
    block_reduce_store( e_i, t_i, A.part_E, A.part_T );
}

void llg_launch_observe( const StepArgs & A, int solver, int n_blocks, hipStream_t st )
{
    hipLaunchKernelGGL( k_observe, dim3( n_blocks ), dim3( LLG_THREADS ), 0, st, A, solver );
}

// Gradient (SoA) and energy partials of a configuration, for getters and parity tests.
__global__ void __launch_bounds__( LLG_THREADS ) k_gradient( StepArgs A, const double * __restrict__ x0,
                                                            double * __restrict__ grad )
{
    const int blk           = blockIdx.x;
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
    block_reduce_store( e_i, 0.0, A.part_E, A.part_T );
}

// Per-image reduction of the partials; convergence flag (Method_Solver.hpp:232-239)
// with the index of the step at which the run converged.
__global__ void k_finalize( const ImageDev * __restrict__ imgs, int n_images, const double * __restrict__ part_E,
                            const double * __restrict__ part_T, double * __restrict__ obs, int * __restrict__ stop,
                            long long * __restrict__ stop_step, long long step_index, double force_convergence,
                            int check_convergence )
{
    __shared__ double sE[256], sT[256];
    if( *stop )
        return;
    double Tall = 0;
    for( int img = 0; img < n_images; ++img )
    {
        const ImageDev & I = imgs[img];
        double e = 0, t = 0;
        for( int k = threadIdx.x; k < I.n_blocks; k += blockDim.x )
        {
            e += part_E[I.block_begin + k];
            t = fmax( t, part_T[I.block_begin + k] );
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
            obs[2 * img]     = sE[0];
            obs[2 * img + 1] = sqrt( sT[0] );
        }
        Tall = fmax( Tall, sT[0] );
        __syncthreads();
    }
    if( threadIdx.x == 0 )
    {
        if( check_convergence && sqrt( Tall ) < force_convergence )
        {
            *stop      = 1;
            *stop_step = step_index;
        }
    }
}

// ---------------------------------------------------------------------------
// host launchers
// ---------------------------------------------------------------------------
template<int SOLVER, int STAGE>
static void launch_stage( const StepArgs & A, int n_blocks, hipStream_t st )
{
    hipLaunchKernelGGL( ( k_llg_stage<SOLVER, STAGE> ), dim3( n_blocks ), dim3( LLG_THREADS ), 0, st, A );
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

void llg_launch_finalize( const StepArgs & A, int n_images, double * obs, long long * stop_step, long long step_index,
                          double force_convergence, int check_convergence, hipStream_t st )
{
    hipLaunchKernelGGL( k_finalize, dim3( 1 ), dim3( 256 ), 0, st, A.imgs, n_images, A.part_E, A.part_T, obs, A.stop,
                        stop_step, step_index, force_convergence, check_convergence );
}

} // namespace spirit_hip
