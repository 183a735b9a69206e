// Per-spin device helpers shared by the LLG and GNEB kernels (FP64): vector algebra with
// the reference's (Eigen's) evaluation order, the counter-based thermal generator, the
// neighbour index, the fused local gradient and the block reductions.
//
// Restated reference code (paths relative to the reference checkout):
//   Hamiltonian_Heisenberg::Gradient_and_Energy  core/src/engine/Hamiltonian_Heisenberg.cpp:704-864
//   Vectormath::idx_from_pair                    core/include/engine/Vectormath.hpp:437-528
//   Vectormath::rotate, sib_transform            core/src/engine/Vectormath.cpp:474-485,
//                                                core/src/engine/Solver_Kernels.cpp:15-48
//   Method_LLG::Calculate_Force_Virtual          core/src/engine/Method_LLG.cpp:132-226
//   Method_LLG::Prepare_Thermal_Field            core/src/engine/Method_LLG.cpp:66-110
#pragma once

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
// a / |a| with one division (a product by the reciprocal norm); zero vectors stay zero
__device__ __forceinline__ void normalize3_rcp( double * a )
{
    double z = dot3( a, a );
    if( z > 0 )
    {
        const double r = 1.0 / sqrt( z );
        a[0] *= r;
        a[1] *= r;
        a[2] *= r;
    }
}
// sin and cos of an angle in [0, 0.25] from their Taylor series (through x^13 and x^14:
// the first omitted terms are below 1e-20 relative), else from sincos.
__device__ __forceinline__ void sincos_small( double t, double * sn, double * cs )
{
    if( t <= 0.25 )
    {
        const double x = t * t;
        double ps      = 1.0 / 6227020800.0;
        ps             = ps * x - 1.0 / 39916800.0;
        ps             = ps * x + 1.0 / 362880.0;
        ps             = ps * x - 1.0 / 5040.0;
        ps             = ps * x + 1.0 / 120.0;
        ps             = ps * x - 1.0 / 6.0;
        *sn            = t + t * ( ps * x );
        double pc      = -1.0 / 87178291200.0;
        pc             = pc * x + 1.0 / 479001600.0;
        pc             = pc * x - 1.0 / 3628800.0;
        pc             = pc * x + 1.0 / 40320.0;
        pc             = pc * x - 1.0 / 720.0;
        pc             = pc * x + 1.0 / 24.0;
        pc             = pc * x - 0.5;
        *cs            = 1.0 + pc * x;
    }
    else
        sincos( t, sn, cs );
}
// Rodrigues rotation (as rotate3) for the solver stages: the angle of one time step is
// small, so its sine and cosine come from sincos_small
__device__ __forceinline__ void rotate3_step( const double * v, const double * axis, double angle, double * out )
{
    double sn, c;
    sincos_small( angle, &sn, &c );
    double cr[3];
    cross3( axis, v, cr );
    double d = dot3( axis, v );
    double o[3];
#pragma unroll
    for( int q = 0; q < 3; ++q )
        o[q] = v[q] * c + cr[q] * sn + axis[q] * d * ( 1 - c );
    out[0] = o[0], out[1] = o[1], out[2] = o[2];
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
    double sn, cs;
    sincos( 2.0 * Pi * u2, &sn, &cs ); // one argument reduction for both
    z[0] = r1 * cs;
    z[1] = r1 * sn;
    z[2]      = r2 * cos( 2.0 * Pi * u4 );
}

// floor(n / f.d) for 0 <= n < 2^31 (FastDiv)
__device__ __forceinline__ int fast_div( int n, const FastDiv & f )
{
    const uint64_t p = (uint64_t)(uint32_t)n * f.m_lo + ( (uint64_t)( (uint32_t)n * f.m_hi ) << 32 );
    return (int)( p >> f.shift );
}

// basis atom of spin i
__device__ __forceinline__ int basis_of( const HamDev & H, int i )
{
    return i - fast_div( i, H.div_nb ) * H.nb;
}

// SoA field access: component q of spin i at f[q * nos + i]
__device__ __forceinline__ void load3( const double * f, int nos, int i, double * v )
{
    v[0] = f[i], v[1] = f[nos + i], v[2] = f[2 * nos + i];
}
__device__ __forceinline__ void store3( double * f, int nos, int i, const double * v )
{
    f[i] = v[0], f[nos + i] = v[1], f[2 * nos + i] = v[2];
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

// neighbours a spin gathers into registers before using them (larger tables loop)
constexpr int LG_MAX_NB = GATHER_PAIRS;

// The exchange neighbours of spin i (tables of up to LG_MAX_NB pairs), all loads issued
// before any is used.  A pair that does not start at this basis atom or leaves an open
// boundary reads the spin itself and yields exactly zero (its loaded values are replaced).
__device__ __forceinline__ void gather_exchange_neighbours( const HamDev & H, const double * __restrict__ x, int nos,
                                                            int i, int ib, int a, int b, int c,
                                                            double ( &se )[LG_MAX_NB][3] )
{
    // away from every boundary a neighbour is the site's index plus a fixed offset
    const bool interior = a >= H.halo[0] && a < H.Na - H.halo[0] && b >= H.halo[1] && b < H.Nb - H.halo[1]
                          && c >= H.halo[2] && c < H.Nc - H.halo[2];
#pragma unroll
    for( int k = 0; k < LG_MAX_NB; ++k )
        if( k < H.n_exch )
        {
            const PairDev & p = H.exch_g[k];
            const int j = p.i != ib ? -1 : interior ? i + p.delta : pair_neighbour( H, a, b, c, p );
            load3( x, nos, j >= 0 ? j : i, se[k] );
            if( j < 0 )
                se[k][0] = se[k][1] = se[k][2] = 0.0;
        }
}

// The quadruplet part of spin i's gradient, gathered: Gradient_Quadruplet
// (Hamiltonian_Heisenberg.cpp:1073-1117) scatters, for every quadruplet q and every cell, the terms
//   i: Q s_j (s_k.s_l),  j: Q s_i (s_k.s_l),  k: Q (s_i.s_j) s_l,  l: Q (s_i.s_j) s_k
// (subtracted from the gradient).  Spin i collects the terms of every (q, role) whose basis atom
// is its own, from the cell the quadruplet is anchored at (its own cell shifted back by the
// role's translation, wrapped along periodic axes); the quadruplet counts only if all four of
// its spins exist from that anchor.  Returns the sum t; the spin's energy share is -t.s/4
// (E_Quadruplet, :513-557).
__device__ __forceinline__ void quadruplet_gradient( const HamDev & H, const double * __restrict__ x, int nos, int ib,
                                                     int a, int b, int c, double * t )
{
    t[0] = t[1] = t[2] = 0;
    const int N[3] = { H.Na, H.Nb, H.Nc };
    for( int k = 0; k < H.n_quad; ++k )
    {
        const QuadDev & q = H.quad[k];
        const int basis[4] = { q.i, q.j, q.k, q.l };
        for( int r = 0; r < 4; ++r )
        {
            if( basis[r] != ib )
                continue;
            const int * d = r == 1 ? q.dj : r == 2 ? q.dk : q.dl;
            int o[3]      = { a, b, c };
            bool ok       = true;
            if( r > 0 )
            {
                const int p[3] = { a, b, c };
                for( int u = 0; u < 3; ++u )
                {
                    int v = p[u] - d[u];
                    if( H.bc[u] )
                    {
                        if( v < 0 )
                            v += N[u];
                        if( v >= N[u] )
                            v -= N[u];
                    }
                    else if( v < 0 || v >= N[u] )
                        ok = false;
                    o[u] = v;
                }
            }
            if( !ok )
                continue;
            const PairDev pj = { q.i, q.j, q.dj[0], q.dj[1], q.dj[2], 0, 0, 0, 0, 0, { 0, 0, 0 } };
            const PairDev pk = { q.i, q.k, q.dk[0], q.dk[1], q.dk[2], 0, 0, 0, 0, 0, { 0, 0, 0 } };
            const PairDev pl = { q.i, q.l, q.dl[0], q.dl[1], q.dl[2], 0, 0, 0, 0, 0, { 0, 0, 0 } };
            const int si_ = q.i + H.nb * ( o[0] + H.Na * ( o[1] + H.Nb * o[2] ) );
            const int sj_ = pair_neighbour( H, o[0], o[1], o[2], pj );
            const int sk_ = pair_neighbour( H, o[0], o[1], o[2], pk );
            const int sl_ = pair_neighbour( H, o[0], o[1], o[2], pl );
            if( sj_ < 0 || sk_ < 0 || sl_ < 0 )
                continue;
            double vi[3], vj[3], vk[3], vl[3];
            load3( x, nos, si_, vi );
            load3( x, nos, sj_, vj );
            load3( x, nos, sk_, vk );
            load3( x, nos, sl_, vl );
            if( r < 2 )
            {
                const double kl = dot3( vk, vl );
                const double * o2 = r == 0 ? vj : vi;
                for( int u = 0; u < 3; ++u )
                    t[u] += q.Q * o2[u] * kl;
            }
            else
            {
                const double ij   = dot3( vi, vj );
                const double * o2 = r == 2 ? vl : vk;
                for( int u = 0; u < 3; ++u )
                    t[u] += q.Q * ij * o2[u];
            }
        }
    }
}

// Gradient_and_Energy for one spin: g (all terms) and its energy share e
// (1/2 g.s over aniso+exchange+DMI+DDI, plus the cubic and Zeeman energies).
__device__ __forceinline__ void local_gradient( const HamDev & H, const double * __restrict__ x, int nos, int i,
                                                const double * __restrict__ hddi, double * g, double & e )
{
    const int cell = fast_div( i, H.div_nb );
    const int ib   = i - cell * H.nb;
    const int rest = fast_div( cell, H.div_Na );
    const int a    = cell - rest * H.Na;
    const int c    = fast_div( rest, H.div_Nb );
    const int b    = rest - c * H.Nb;
    const double s[3] = { x[i], x[nos + i], x[2 * nos + i] };
    // Every load this spin needs is issued before the first use: the per-site K_i, the DDI
    // field and the neighbour spins (up to LG_MAX_NB of them gathered into registers), so
    // the memory latencies overlap instead of adding up pair after pair.
    const double Ki = H.has_site_K ? H.site_K[i] : 0.0;
    double hd[3]    = { 0, 0, 0 };
    if( hddi )
        load3( hddi, nos, i, hd );
    const bool gather = H.n_exch <= LG_MAX_NB;
    double se[LG_MAX_NB][3];
    if( gather )
        gather_exchange_neighbours( H, x, nos, i, ib, a, b, c, se );
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
            g[q] -= 2.0 * Ki * H.site_n[q] * d;
    }
    // exchange, :822-842 (gather over the redundant table)
    if( gather )
    {
#pragma unroll
        for( int k = 0; k < LG_MAX_NB; ++k )
            if( k < H.n_exch )
            {
                const double J = H.exch_g[k].J;
#pragma unroll
                for( int q = 0; q < 3; ++q )
                    g[q] -= J * se[k][q];
            }
    }
    else
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
    // DMI, :844-864 (on the exchange pairs' neighbours when the tables coincide)
    if( gather && H.dmi_same )
    {
#pragma unroll
        for( int k = 0; k < LG_MAX_NB; ++k )
            if( k < H.n_dmi )
            {
                const PairDev & p = H.dmi_g[k];
                double cr[3];
                cross3( se[k], p.n, cr );
#pragma unroll
                for( int q = 0; q < 3; ++q )
                    g[q] -= p.D * cr[q];
            }
    }
    else
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
        g[0] -= hd[0];
        g[1] -= hd[1];
        g[2] -= hd[2];
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
    // quadruplets last, :753-765
    if( H.quad_in_step )
    {
        double t[3];
        quadruplet_gradient( H, x, nos, ib, a, b, c, t );
        g[0] -= t[0];
        g[1] -= t[1];
        g[2] -= t[2];
        e += -0.25 * dot3( t, s );
    }
}

// Energy contributions of spin i (Hamiltonian_Heisenberg::E_Zeeman, E_Anisotropy,
// E_Cubic_Anisotropy, E_Exchange, E_DMI, E_DDI_FFT, :306-535), in the order
// ENERGY_TERMS lists; the pair terms count half of each pair at each of its spins.
constexpr int N_ENERGY_TERMS = 7; // Zeeman, Anisotropy, Cubic anisotropy, Exchange, DMI, DDI, Quadruplets
__device__ __forceinline__ void local_energy_terms( const HamDev & H, const double * __restrict__ x, int nos, int i,
                                                    const double * __restrict__ hddi, double * E )
{
    const int cell = fast_div( i, H.div_nb );
    const int ib   = i - cell * H.nb;
    const int rest = fast_div( cell, H.div_Na );
    const int a    = cell - rest * H.Na;
    const int c    = fast_div( rest, H.div_Nb );
    const int b    = rest - c * H.Nb;
    const double s[3] = { x[i], x[nos + i], x[2 * nos + i] };
    for( int t = 0; t < N_ENERGY_TERMS; ++t )
        E[t] = 0;
    if( H.has_zeeman )
        E[0] -= H.mu[ib] * H.B * dot3( H.Bn, s );
    for( int k = 0; k < H.n_aniso; ++k )
        if( H.aniso_idx[k] == ib )
        {
            const double d = dot3( H.aniso_n[k], s );
            E[1] -= H.aniso_K[k] * pow( d, 2.0 );
        }
    if( H.has_site_K )
    {
        const double d = dot3( H.site_n, s );
        E[1] -= H.site_K[i] * pow( d, 2.0 );
    }
    for( int k = 0; k < H.n_cubic; ++k )
        if( H.cubic_idx[k] == ib )
            E[2] -= H.cubic_K[k] / 2 * ( pow( s[0], 4.0 ) + pow( s[1], 4.0 ) + pow( s[2], 4.0 ) );
    for( int k = 0; k < H.n_exch; ++k )
    {
        const PairDev & p = H.exch[k];
        if( p.i != ib )
            continue;
        const int j = pair_neighbour( H, a, b, c, p );
        if( j >= 0 )
        {
            const double sj[3] = { x[j], x[nos + j], x[2 * nos + j] };
            E[3] -= 0.5 * p.J * dot3( s, sj );
        }
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
            E[4] -= 0.5 * p.D * dot3( p.n, cr );
        }
    }
    if( hddi )
    {
        // the DDI gradient is -h (h = mu_i (D * mu s)_i / M): E_i = 1/2 s . (-h)
        const double g[3] = { -hddi[i], -hddi[nos + i], -hddi[2 * nos + i] };
        E[5] += 0.5 * dot3( s, g );
    }
    // E_Quadruplet, :513-557: each of the four spins takes -Q (s_i.s_j)(s_k.s_l) / 4
    if( H.n_quad > 0 )
    {
        double t[3];
        quadruplet_gradient( H, x, nos, ib, a, b, c, t );
        E[6] += -0.25 * dot3( t, s );
    }
}

// The derivative of the spin field along the current direction j at spin i:
// Vectormath::jacobian (Vectormath.cpp:816-898) times j (Method_LLG.cpp:188-196).  Along
// each lattice translation t the central difference (s(+t) - s(-t)) / 2, one-sided (factor 1)
// where an open boundary removes a neighbour; J j = sum_t diff_t w_t with w = B^-1 j.
__device__ __forceinline__ void stt_direction_gradient( const HamDev & H, const LLGDev & P,
                                                        const double * __restrict__ x, int nos, int i, double * g )
{
    const int cell = fast_div( i, H.div_nb );
    const int ib   = i - cell * H.nb;
    const int rest = fast_div( cell, H.div_Na );
    const int a    = cell - rest * H.Na;
    const int c    = fast_div( rest, H.div_Nb );
    const int b    = rest - c * H.Nb;
    const int pos[3] = { a, b, c };
    const int n[3]   = { H.Na, H.Nb, H.Nc };
    g[0] = g[1] = g[2] = 0;
#pragma unroll
    for( int t = 0; t < 3; ++t )
    {
        // neighbour coordinate along t at +1 and -1 (linear_idx, Vectormath.cpp:773-814)
        int up = pos[t] + 1, dn = pos[t] - 1;
        bool has_up = up < n[t], has_dn = dn >= 0;
        if( !has_up && H.bc[t] )
            up -= n[t], has_up = true;
        if( !has_dn && H.bc[t] )
            dn += n[t], has_dn = true;
        double factor = 0.5;
        int q_up[3] = { a, b, c }, q_dn[3] = { a, b, c };
        q_up[t] = up, q_dn[t] = dn;
        const int j_up = has_up ? ib + H.nb * ( q_up[0] + H.Na * ( q_up[1] + H.Nb * q_up[2] ) ) : i;
        const int j_dn = has_dn ? ib + H.nb * ( q_dn[0] + H.Na * ( q_dn[1] + H.Nb * q_dn[2] ) ) : i;
        if( !has_up )
            factor *= 2;
        if( !has_dn )
            factor *= 2;
        double m0[3], m1[3];
        load3( x, nos, j_up, m0 );
        load3( x, nos, j_dn, m1 );
#pragma unroll
        for( int q = 0; q < 3; ++q )
            g[q] += factor * ( m0[q] - m1[q] ) * P.stt_w[t];
    }
}

// Method_LLG::Calculate_Force_Virtual, Method_LLG.cpp:132-226
// inv_mu = 1 / mu_s of the spin (the dynamics' division by mu_s, as a product); sg is the
// spin's STT direction derivative (stt_direction_gradient) when P.stt == STT_GRADIENT
__device__ __forceinline__ void virtual_force( const LLGDev & P, int solver, double inv_mu, const double * s,
                                               const double * F, const double * xi, double * fv,
                                               const double * sg = nullptr )
{
    const double a = P.damping;
    if( solver == SOLVER_LBFGS_OSO || solver == SOLVER_LBFGS_ATLAS || P.direct_minimization || solver == SOLVER_VP
        || solver == SOLVER_VP_OSO )
    {
        // Method_LLG.cpp:160-171: s x F, scaled by dt gamma / mu_B except for the L-BFGS solvers
        const double dtg = ( solver == SOLVER_LBFGS_OSO || solver == SOLVER_LBFGS_ATLAS ) ? 1.0 : P.dt * gamma_e / mu_B;
        double c[3];
        cross3( s, F, c );
#pragma unroll
        for( int q = 0; q < 3; ++q )
            fv[q] = dtg * c[q];
        return;
    }
    const double dtg = P.vf_dtg; // dt gamma / mu_B / (1 + a^2), Method_LLG.cpp:149
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
        v[q] *= inv_mu;
    if( P.stt != STT_OFF )
    {
        // monolayer (:203-208): c1 = -dtg a_j (alpha - beta), c2 = -dtg a_j (1 + beta alpha);
        // gradient (:184-201): the same with the opposite signs and p = (j . grad) s
        const double * p = P.stt == STT_GRADIENT ? sg : P.stt_p;
        cross3( p, s, c );
#pragma unroll
        for( int q = 0; q < 3; ++q )
            v[q] += P.stt_c1 * p[q];
#pragma unroll
        for( int q = 0; q < 3; ++q )
            v[q] += P.stt_c2 * c[q];
    }
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
    // a temperature gradient replaces T by the site's T_i (Method_LLG.cpp:80-93)
    const double T   = P.temp_site ? P.temp_site[i] : P.temperature;
    const double f   = eps * sqrt( T / mu );
#pragma unroll
    for( int q = 0; q < 3; ++q )
        xi[q] = f * z[q];
}

// The hardware deals consecutive blocks of a launch to the 8 XCDs round-robin.  Renumber
// them so that each XCD runs one contiguous range of logical blocks: the lattice rows and
// planes a block's neighbours live in are then read through the same L2.  A bijection on
// [0, gridDim.x) (a tail of fewer than 8 blocks keeps its numbers).
__device__ __forceinline__ int xcd_contiguous_block()
{
    const int nblk = gridDim.x;
    const int full = ( nblk / 8 ) * 8;
    const int b    = blockIdx.x;
    return b < full ? ( b % 8 ) * ( full / 8 ) + b / 8 : b;
}

// Order in which the blocks of one image visit the lattice in the stage kernels: rows of
// cells along a (all basis atoms, whole blocks per row) in tiles of 2^row_tile_log2
// consecutive b, each tile swept through every c plane, so the rows b +- 1 and c +- 1 a
// block reads were fetched into the same L2 by blocks that ran just before it instead of
// a whole 2D plane earlier.  A bijection on the image's blocks; the identity where the
// rows do not hold whole blocks, Nb is not a multiple of the tile or the lattice is 2D.
__device__ __forceinline__ int stage_block_order( const HamDev & H, int q )
{
    const int R    = H.nb * H.Na;
    const int TL   = H.row_tile_log2;
    const int TILE = 1 << TL;
    if( TL <= 0 || R % LLG_THREADS != 0 || ( H.Nb & ( TILE - 1 ) ) != 0 || H.Nc == 1 )
        return q;
    const int per_row = R / LLG_THREADS;
    const int j       = q % per_row;
    int t             = q / per_row;
    const int bi      = t & ( TILE - 1 );
    t >>= TL;
    const int c  = t % H.Nc;
    const int bt = t / H.Nc;
    return ( ( bt << TL ) + bi + H.Nb * c ) * per_row + j;
}

// block-wide reduction of (sum e, max t) into partials[blk]
__device__ __forceinline__ void block_reduce_store( double e, double t, double * part_E, double * part_T, int blk )
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
        part_E[blk] = E;
        part_T[blk] = T;
    }
}



} // namespace spirit_hip
