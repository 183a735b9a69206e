// Mixed-radix Stockham FFT on lines held in LDS (FP64 complex), for gfx950.
//
// Lines of length n live in LDS at lds[line * ls + i].  Every stage reads all of
// its butterfly inputs into registers, waits at a barrier, then writes the
// outputs in place at their Stockham (self-sorting) positions, so one LDS copy
// of the data is enough.  A thread holds at most 16 complex values per stage
// (16/R butterflies of radix R), which bounds nlines * n by 16 * blockDim.
//
// Stage s of radix R with Ns = product of the previous radices:
//   inputs   x[j + r*n/R], r < R, j < n/R
//   twiddle  x_r *= w_n^(r * (j mod Ns) * n/(Ns*R))
//   DFT_R, outputs to (j/Ns)*Ns*R + (j mod Ns) + k*Ns.
// Forward transforms use w_n = exp(-2 pi i/n); inverse ones the conjugate.
// Both are unnormalised, like the reference's FFT backends.
#pragma once

#include "device_types.hpp"

namespace spirit_hip
{

__device__ __forceinline__ double2 cmul( double2 a, double2 b )
{
    return make_double2( a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x );
}
__device__ __forceinline__ double2 cadd( double2 a, double2 b )
{
    return make_double2( a.x + b.x, a.y + b.y );
}
__device__ __forceinline__ double2 csub( double2 a, double2 b )
{
    return make_double2( a.x - b.x, a.y - b.y );
}
__device__ __forceinline__ double2 cconj( double2 a )
{
    return make_double2( a.x, -a.y );
}
// multiply by -i (forward) or +i (inverse)
template<bool INV>
__device__ __forceinline__ double2 mul_mi( double2 a )
{
    return INV ? make_double2( -a.y, a.x ) : make_double2( a.y, -a.x );
}

template<int R, bool INV>
__device__ __forceinline__ void dft_small( double2 * v, const double2 * W, int n )
{
    if constexpr( R == 2 )
    {
        double2 a = v[0], b = v[1];
        v[0]      = cadd( a, b );
        v[1]      = csub( a, b );
    }
    else if constexpr( R == 4 )
    {
        double2 a0 = cadd( v[0], v[2] ), a1 = csub( v[0], v[2] );
        double2 b0 = cadd( v[1], v[3] ), b1 = mul_mi<INV>( csub( v[1], v[3] ) );
        v[0]       = cadd( a0, b0 );
        v[2]       = csub( a0, b0 );
        v[1]       = cadd( a1, b1 );
        v[3]       = csub( a1, b1 );
    }
    else if constexpr( R == 8 )
    {
        // DIT: even/odd radix-4 halves, then the w8^k combination.
        double2 e[4] = { v[0], v[2], v[4], v[6] };
        double2 o[4] = { v[1], v[3], v[5], v[7] };
        dft_small<4, INV>( e, W, n );
        dft_small<4, INV>( o, W, n );
        const double h = 0.70710678118654752440084436210484903928483593768847;
        // w8^1 = (1 -+ i)/sqrt2, w8^2 = -+i, w8^3 = (-1 -+ i)/sqrt2
        double2 t1 = INV ? make_double2( h * ( o[1].x - o[1].y ), h * ( o[1].x + o[1].y ) )
                         : make_double2( h * ( o[1].x + o[1].y ), h * ( o[1].y - o[1].x ) );
        double2 t2 = mul_mi<INV>( o[2] );
        double2 t3 = INV ? make_double2( -h * ( o[3].x + o[3].y ), h * ( o[3].x - o[3].y ) )
                         : make_double2( h * ( o[3].y - o[3].x ), -h * ( o[3].x + o[3].y ) );
        v[0] = cadd( e[0], o[0] );
        v[4] = csub( e[0], o[0] );
        v[1] = cadd( e[1], t1 );
        v[5] = csub( e[1], t1 );
        v[2] = cadd( e[2], t2 );
        v[6] = csub( e[2], t2 );
        v[3] = cadd( e[3], t3 );
        v[7] = csub( e[3], t3 );
    }
    else
    {
        // generic odd radix from the n-th roots: exp(-2 pi i m/R) = W[m * n/R]
        double2 out[R];
        const int step = n / R;
#pragma unroll
        for( int k = 0; k < R; ++k )
        {
            double2 acc = v[0];
#pragma unroll
            for( int r = 1; r < R; ++r )
            {
                double2 w = W[( ( r * k ) % R ) * step];
                if( INV )
                    w.y = -w.y;
                acc = cadd( acc, cmul( v[r], w ) );
            }
            out[k] = acc;
        }
#pragma unroll
        for( int k = 0; k < R; ++k )
            v[k] = out[k];
    }
}

template<int R, bool INV>
__device__ __forceinline__ void fft_stage( double2 * lds, int ls, int nlines, int n, int Ns, const double2 * W )
{
    constexpr int PER  = 16 / R;
    const int nbf_line = n / R;
    const int nbf      = nlines * nbf_line;
    double2 v[PER][R];
#pragma unroll
    for( int k = 0; k < PER; ++k )
    {
        int t = threadIdx.x + k * blockDim.x;
        if( t < nbf )
        {
            int line               = t / nbf_line;
            int j                  = t - line * nbf_line;
            const double2 * base   = lds + line * ls;
#pragma unroll
            for( int r = 0; r < R; ++r )
                v[k][r] = base[j + r * nbf_line];
        }
    }
    __syncthreads();
#pragma unroll
    for( int k = 0; k < PER; ++k )
    {
        int t = threadIdx.x + k * blockDim.x;
        if( t < nbf )
        {
            int line        = t / nbf_line;
            int j           = t - line * nbf_line;
            int jm          = j % Ns;
            int tstep       = jm * ( n / ( Ns * R ) );
            double2 * base  = lds + line * ls;
#pragma unroll
            for( int r = 1; r < R; ++r )
            {
                double2 w = W[r * tstep];
                if( INV )
                    w.y = -w.y;
                v[k][r] = cmul( v[k][r], w );
            }
            dft_small<R, INV>( v[k], W, n );
            int idxD = ( j / Ns ) * Ns * R + jm;
#pragma unroll
            for( int r = 0; r < R; ++r )
                base[idxD + r * Ns] = v[k][r];
        }
    }
    __syncthreads();
}

// In-place FFT of nlines lines (stride ls) of length L.n held in LDS.
// Requires nlines * L.n <= 16 * blockDim (checked on the host) and a barrier
// between the caller's LDS writes and this call.
template<bool INV>
__device__ void lds_fft( double2 * lds, int ls, int nlines, const FftLen & L )
{
    int Ns = 1;
    for( int s = 0; s < L.n_stages; ++s )
    {
        const int R = L.radix[s];
        switch( R )
        {
            case 2: fft_stage<2, INV>( lds, ls, nlines, L.n, Ns, L.W ); break;
            case 3: fft_stage<3, INV>( lds, ls, nlines, L.n, Ns, L.W ); break;
            case 4: fft_stage<4, INV>( lds, ls, nlines, L.n, Ns, L.W ); break;
            case 5: fft_stage<5, INV>( lds, ls, nlines, L.n, Ns, L.W ); break;
            case 7: fft_stage<7, INV>( lds, ls, nlines, L.n, Ns, L.W ); break;
            case 8: fft_stage<8, INV>( lds, ls, nlines, L.n, Ns, L.W ); break;
            case 11: fft_stage<11, INV>( lds, ls, nlines, L.n, Ns, L.W ); break;
            case 13: fft_stage<13, INV>( lds, ls, nlines, L.n, Ns, L.W ); break;
            default: break;
        }
        Ns *= R;
    }
}

} // namespace spirit_hip
