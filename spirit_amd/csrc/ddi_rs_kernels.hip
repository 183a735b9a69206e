// Dipole-dipole field in real space (FP64), hand-written for gfx950: the reference's two
// non-FFT methods, both selected by ddi_method "cutoff" (Hamiltonian_Heisenberg.cpp:866-878):
//   radius >= 0: Gradient_DDI_Cutoff (Hamiltonian_Heisenberg.cpp:880-918), a sum over the
//                pairs within the radius (Neighbours.cpp:158-246) with the lattice's
//                boundary conditions;
//   radius <  0: Gradient_DDI_Direct (Hamiltonian_Heisenberg.cpp:1016-1071), the all-pairs
//                sum over the lattice positions with periodic images along periodic axes.
// Both write h_i (SoA) with the gradient's dipolar part g_i = -h_i, the same field the FFT
// convolution writes, so every force kernel consumes it unchanged.
#include "ddi_rs.hpp"

namespace spirit_hip
{

namespace
{
constexpr int RS_THREADS = 256;

__device__ __forceinline__ int rs_div( int n, const FastDiv & f )
{
    const uint64_t p = (uint64_t)(uint32_t)n * f.m_lo + ( (uint64_t)( (uint32_t)n * f.m_hi ) << 32 );
    return (int)( p >> f.shift );
}

// neighbour of cell (a, b, c) along translation (da, db, dc) to basis atom jb with the
// lattice's boundary conditions (Vectormath.hpp:437-528); -1 outside an open boundary
__device__ __forceinline__ int rs_neighbour( const DdiRsDev & D, int a, int b, int c, const DdiPairDev & p )
{
    int t[3]       = { a + p.da, b + p.db, c + p.dc };
    const int n[3] = { D.Na, D.Nb, D.Nc };
    const int d[3] = { p.da, p.db, p.dc };
#pragma unroll
    for( int k = 0; k < 3; ++k )
    {
        if( abs( d[k] ) > n[k] )
            return -1;
        if( D.bc[k] || ( 0 <= t[k] && t[k] < n[k] ) )
        {
            if( t[k] < 0 )
                t[k] += n[k];
            if( t[k] >= n[k] )
                t[k] -= n[k];
        }
        else
            return -1;
    }
    return p.j + D.nb * ( t[0] + D.Na * ( t[1] + D.Nb * t[2] ) );
}
} // namespace

// One thread per spin, the cutoff pairs in the reference's pair order.
__global__ void __launch_bounds__( RS_THREADS ) k_ddi_cutoff( DdiRsDev D, const double * __restrict__ s,
                                                              double * __restrict__ h )
{
    const int i = blockIdx.x * RS_THREADS + threadIdx.x;
    if( i >= D.nos )
        return;
    const int nos  = D.nos;
    const int cell = rs_div( i, D.div_nb );
    const int ib   = i - cell * D.nb;
    const int rest = rs_div( cell, D.div_Na );
    const int a    = cell - rest * D.Na;
    const int c    = rs_div( rest, D.div_Nb );
    const int b    = rest - c * D.Nb;
    const double mui = D.mu[ib];
    double h0 = 0, h1 = 0, h2 = 0;
    for( int k = 0; k < D.n_pairs; ++k )
    {
        const DdiPairDev & p = D.pairs[k];
        if( p.i != ib )
            continue;
        const int j = rs_neighbour( D, a, b, c, p );
        if( j < 0 )
            continue;
        const double sj[3] = { s[j], s[nos + j], s[2 * nos + j] };
        const double sn    = sj[0] * p.n[0] + sj[1] * p.n[1] + sj[2] * p.n[2];
        const double f     = D.mu[p.j] * mui * p.c;
        h0 += f * ( 3 * p.n[0] * sn - sj[0] );
        h1 += f * ( 3 * p.n[1] * sn - sj[1] );
        h2 += f * ( 3 * p.n[2] * sn - sj[2] );
    }
    h[i]           = h0;
    h[nos + i]     = h1;
    h[2 * nos + i] = h2;
}

// One thread per spin i; the partner spins j pass through LDS in tiles of RS_THREADS (all
// threads read the same entry: broadcasts), summed in index order as the reference does.
__global__ void __launch_bounds__( RS_THREADS ) k_ddi_direct( DdiRsDev D, const double * __restrict__ s,
                                                              double * __restrict__ h )
{
    __shared__ double tile[7][RS_THREADS];
    const int nos  = D.nos;
    const int i    = blockIdx.x * RS_THREADS + threadIdx.x;
    const bool act = i < nos;
    const int ii   = act ? i : 0;
    const double pi0 = D.pos[ii], pi1 = D.pos[nos + ii], pi2 = D.pos[2 * nos + ii];
    const double mui = D.mu[ii - rs_div( ii, D.div_nb ) * D.nb];
    const double mult = D.mult;
    double h0 = 0, h1 = 0, h2 = 0;
    for( int j0 = 0; j0 < nos; j0 += RS_THREADS )
    {
        __syncthreads();
        const int j = j0 + threadIdx.x;
        if( j < nos )
        {
            tile[0][threadIdx.x] = D.pos[j];
            tile[1][threadIdx.x] = D.pos[nos + j];
            tile[2][threadIdx.x] = D.pos[2 * nos + j];
            tile[3][threadIdx.x] = s[j];
            tile[4][threadIdx.x] = s[nos + j];
            tile[5][threadIdx.x] = s[2 * nos + j];
            tile[6][threadIdx.x] = D.mu[j - rs_div( j, D.div_nb ) * D.nb];
        }
        __syncthreads();
        const int nj = min( RS_THREADS, nos - j0 );
        for( int jj = 0; jj < nj; ++jj )
        {
            const double d0 = tile[0][jj] - pi0, d1 = tile[1][jj] - pi1, d2 = tile[2][jj] - pi2;
            double Dxx = 0, Dxy = 0, Dxz = 0, Dyy = 0, Dyz = 0, Dzz = 0;
            for( int ia = -D.img[0]; ia <= D.img[0]; ++ia )
                for( int ib = -D.img[1]; ib <= D.img[1]; ++ib )
                    for( int ic = -D.img[2]; ic <= D.img[2]; ++ic )
                    {
                        // diff + (ia*Na) a_0 lc + (ib*Nb) a_1 lc + (ic*Nc) a_2 lc, added in that order
                        const double fa = ia * D.Na, fb = ib * D.Nb, fc = ic * D.Nc;
                        const double x = ( ( d0 + fa * D.bv[0][0] * D.lc ) + fb * D.bv[1][0] * D.lc ) + fc * D.bv[2][0] * D.lc;
                        const double y = ( ( d1 + fa * D.bv[0][1] * D.lc ) + fb * D.bv[1][1] * D.lc ) + fc * D.bv[2][1] * D.lc;
                        const double z = ( ( d2 + fa * D.bv[0][2] * D.lc ) + fb * D.bv[1][2] * D.lc ) + fc * D.bv[2][2] * D.lc;
                        const double d = sqrt( x * x + y * y + z * z );
                        if( d > 1e-10 )
                        {
                            const double d3 = d * d * d;
                            const double d5 = d * d * d * d * d;
                            Dxx += mult * ( 3 * x * x / d5 - 1 / d3 );
                            Dxy += mult * 3 * x * y / d5;
                            Dxz += mult * 3 * x * z / d5;
                            Dyy += mult * ( 3 * y * y / d5 - 1 / d3 );
                            Dyz += mult * 3 * y * z / d5;
                            Dzz += mult * ( 3 * z * z / d5 - 1 / d3 );
                        }
                    }
            const double m0 = tile[3][jj], m1 = tile[4][jj], m2 = tile[5][jj], muj = tile[6][jj];
            h0 += ( Dxx * m0 + Dxy * m1 + Dxz * m2 ) * mui * muj;
            h1 += ( Dxy * m0 + Dyy * m1 + Dyz * m2 ) * mui * muj;
            h2 += ( Dxz * m0 + Dyz * m1 + Dzz * m2 ) * mui * muj;
        }
    }
    if( act )
    {
        h[i]           = h0;
        h[nos + i]     = h1;
        h[2 * nos + i] = h2;
    }
}

void ddi_rs_launch_field( const DdiRsDev & D, const double * spins, double * h, hipStream_t st )
{
    const int blocks = ( D.nos + RS_THREADS - 1 ) / RS_THREADS;
    if( D.method == DDI_RS_DIRECT )
        hipLaunchKernelGGL( k_ddi_direct, dim3( blocks ), dim3( RS_THREADS ), 0, st, D, spins, h );
    else
        hipLaunchKernelGGL( k_ddi_cutoff, dim3( blocks ), dim3( RS_THREADS ), 0, st, D, spins, h );
}

} // namespace spirit_hip
