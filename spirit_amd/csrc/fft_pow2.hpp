// Register-resident Stockham FFT for power-of-two lengths N >= 8 (FP64 complex).
//
// A line of N points is owned by T = N/8 threads; thread t holds x[t + T*k],
// k = 0..7, in registers.  Stages are radix 8 (a final radix 2 or 4 when log2 N
// is not a multiple of 3); a radix-R stage is 8/R butterflies per thread, so the
// register set is always the same eight values.  Between stages the values pass
// through LDS at their Stockham output positions (index padded by one slot every
// eight, which spreads the strided writes over the banks).  Sizes are
// compile-time constants so all index arithmetic folds to shifts.
#pragma once

#include "fft_lds.hpp"

namespace spirit_hip
{

constexpr int ilog2c( int n )
{
    return n <= 1 ? 0 : 1 + ilog2c( n / 2 );
}
constexpr int ipow8c( int s )
{
    return s == 0 ? 1 : 8 * ipow8c( s - 1 );
}
__device__ __forceinline__ constexpr int pad8( int i )
{
    return i + ( i >> 3 );
}

constexpr int ntw_sum( int n8, int rem )
{
    int total = 0, Ns = 8;
    const int nst = n8 + ( rem ? 1 : 0 );
    for( int s = 1; s < nst; ++s, Ns *= 8 )
        total += Ns * ( ( s < n8 ? 8 : ( rem == 1 ? 2 : 4 ) ) - 1 );
    return total > 0 ? total : 1;
}

template<int N>
struct P2
{
    static_assert( N >= 8 && ( N & ( N - 1 ) ) == 0, "power of two >= 8" );
    static constexpr int T   = N / 8;
    static constexpr int LOG = ilog2c( N );
    static constexpr int N8  = LOG / 3;
    static constexpr int REM = LOG % 3;
    static constexpr int NST = N8 + ( REM ? 1 : 0 );
    static constexpr int LS  = N + N / 8 + 1; // padded LDS line stride (complex values)
    // entries of the stage-twiddle table: sum over stages s >= 1 of 8^s * (R_s - 1)
    static constexpr int NTW = ntw_sum( N8, REM );
};

template<int N, int S>
struct P2Stage
{
    static constexpr int R     = ( S < P2<N>::N8 ) ? 8 : ( P2<N>::REM == 1 ? 2 : 4 );
    static constexpr int Ns    = ipow8c( S );
    static constexpr int TWOFF = ipow8c( S ) - 8; // valid for S >= 1
};

// Runs stages S.. of the transform.  On entry v holds this thread's inputs of stage S.
// KEEP = false: on exit the transform is in L[pad8(i)] in natural order (block-synchronised).
// KEEP = true:  the last stage leaves its outputs in registers, v[k] = X[t + T*k] (the
//               same layout as the inputs), and L is free again for the next transform
//               once every thread has passed the barrier that precedes the last stage.
template<int N, bool INV, int S, bool KEEP>
__device__ __forceinline__ void p2_stage( double2 ( &v )[8], double2 * __restrict__ L,
                                          const double2 * __restrict__ tw, int t )
{
    using St           = P2Stage<N, S>;
    constexpr int R    = St::R;
    constexpr int Ns   = St::Ns;
    constexpr int M    = 8 / R;
    constexpr int T    = P2<N>::T;
    constexpr bool LST = ( S + 1 == P2<N>::NST );
    double2 out[8];
#pragma unroll
    for( int m = 0; m < M; ++m )
    {
        const int j  = t + T * m;
        const int jm = j & ( Ns - 1 );
        double2 u[R];
#pragma unroll
        for( int r = 0; r < R; ++r )
            u[r] = v[m + M * r];
        if constexpr( S > 0 )
        {
#pragma unroll
            for( int r = 1; r < R; ++r )
            {
                double2 w = tw[St::TWOFF + jm * ( R - 1 ) + ( r - 1 )];
                if( INV )
                    w.y = -w.y;
                u[r] = cmul( u[r], w );
            }
        }
        dft_small<R, INV>( u, nullptr, N );
        if constexpr( LST && KEEP )
        {
            // last stage: Ns = N/R, so idxD + r*Ns = j + r*N/R = t + T*(m + M*r)
#pragma unroll
            for( int r = 0; r < R; ++r )
                out[m + M * r] = u[r];
        }
        else
        {
            const int idxD = ( j / Ns ) * Ns * R + jm;
#pragma unroll
            for( int r = 0; r < R; ++r )
                L[pad8( idxD + r * Ns )] = u[r];
        }
    }
    if constexpr( LST && KEEP )
    {
#pragma unroll
        for( int k = 0; k < 8; ++k )
            v[k] = out[k];
        return;
    }
    else
    {
        __syncthreads();
        if constexpr( !LST )
        {
#pragma unroll
            for( int k = 0; k < 8; ++k )
                v[k] = L[pad8( t + T * k )];
            __syncthreads();
            p2_stage<N, INV, S + 1, KEEP>( v, L, tw, t );
        }
    }
}

// As p2_stage with KEEP = true, but the values cross LDS real parts first, then imaginary
// parts: a line needs pad8(N) doubles of LDS instead of complex values (twice the blocks
// per CU where LDS limits the occupancy) for two more barriers per stage.  Same arithmetic,
// same result.
// Synchronisation between the LDS writes and reads of an exchange: the whole block, or only
// the wavefront when every line lives inside one wavefront (N/8 <= 64 threads per line and
// lines aligned to wavefronts), which spares the other wavefronts of the block the wait.
template<bool WAVE>
__device__ __forceinline__ void p2_sync()
{
    if constexpr( WAVE )
    {
        __builtin_amdgcn_fence( __ATOMIC_RELEASE, "wavefront" );
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence( __ATOMIC_ACQUIRE, "wavefront" );
    }
    else
        __syncthreads();
}

template<int N, bool INV, int S, bool WAVE = false>
__device__ __forceinline__ void p2_stage_split( double2 ( &v )[8], double * __restrict__ Ld,
                                                const double2 * __restrict__ tw, int t )
{
    using St           = P2Stage<N, S>;
    constexpr int R    = St::R;
    constexpr int Ns   = St::Ns;
    constexpr int M    = 8 / R;
    constexpr int T    = P2<N>::T;
    constexpr bool LST = ( S + 1 == P2<N>::NST );
    double2 out[8];
#pragma unroll
    for( int m = 0; m < M; ++m )
    {
        const int j  = t + T * m;
        const int jm = j & ( Ns - 1 );
        double2 u[R];
#pragma unroll
        for( int r = 0; r < R; ++r )
            u[r] = v[m + M * r];
        if constexpr( S > 0 )
        {
#pragma unroll
            for( int r = 1; r < R; ++r )
            {
                double2 w = tw[St::TWOFF + jm * ( R - 1 ) + ( r - 1 )];
                if( INV )
                    w.y = -w.y;
                u[r] = cmul( u[r], w );
            }
        }
        dft_small<R, INV>( u, nullptr, N );
#pragma unroll
        for( int r = 0; r < R; ++r )
            out[m + M * r] = u[r];
    }
    if constexpr( LST )
    {
        // Ns = N/R: out[m + M*r] is X[t + T*(m + M*r)]
#pragma unroll
        for( int k = 0; k < 8; ++k )
            v[k] = out[k];
    }
    else
    {
#pragma unroll
        for( int m = 0; m < M; ++m )
        {
            const int j    = t + T * m;
            const int idxD = ( j / Ns ) * Ns * R + ( j & ( Ns - 1 ) );
#pragma unroll
            for( int r = 0; r < R; ++r )
                Ld[pad8( idxD + r * Ns )] = out[m + M * r].x;
        }
        p2_sync<WAVE>();
#pragma unroll
        for( int k = 0; k < 8; ++k )
            v[k].x = Ld[pad8( t + T * k )];
        p2_sync<WAVE>();
#pragma unroll
        for( int m = 0; m < M; ++m )
        {
            const int j    = t + T * m;
            const int idxD = ( j / Ns ) * Ns * R + ( j & ( Ns - 1 ) );
#pragma unroll
            for( int r = 0; r < R; ++r )
                Ld[pad8( idxD + r * Ns )] = out[m + M * r].y;
        }
        p2_sync<WAVE>();
#pragma unroll
        for( int k = 0; k < 8; ++k )
            v[k].y = Ld[pad8( t + T * k )];
        p2_sync<WAVE>();
        p2_stage_split<N, INV, S + 1, WAVE>( v, Ld, tw, t );
    }
}

// Transform ending in registers (v[k] = X[t + T*k]) through half-width LDS exchanges.
template<int N, bool INV, bool WAVE = false>
__device__ __forceinline__ void p2_fft_reg_split( double2 ( &v )[8], double * __restrict__ Ld,
                                                  const double2 * __restrict__ tw, int t )
{
    p2_stage_split<N, INV, 0, WAVE>( v, Ld, tw, t );
}

// ---------------------------------------------------------------------------
// Radix-4 variant: a line of N points owned by N/4 threads, thread t holding x[t + (N/4) k],
// k = 0..3; radix-4 stages and a final radix-2 stage when log2 N is odd.  Half the registers
// per line of the radix-8 form (for kernels that hold several lines per thread), one more
// LDS exchange per transform.  Twiddles w_N^e come from the table W of N-th roots.
// ---------------------------------------------------------------------------
constexpr int ipow4c( int s )
{
    return s == 0 ? 1 : 4 * ipow4c( s - 1 );
}

template<int N>
struct P4
{
    static_assert( N >= 4 && ( N & ( N - 1 ) ) == 0, "power of two >= 4" );
    static constexpr int T   = N / 4;
    static constexpr int LOG = ilog2c( N );
    static constexpr int N4  = LOG / 2;
    static constexpr int NST = N4 + ( LOG % 2 );
    static constexpr int LS  = N + N / 8 + 1; // padded LDS line stride (complex values)
};

// Stage S.. of the transform, outputs of the last stage in registers: v[k] = X[t + T*k].
// Every thread of the block must call it; L may be reused once it returns.
template<int N, bool INV, int S>
__device__ __forceinline__ void p4_stage( double2 ( &v )[4], double2 * __restrict__ L,
                                          const double2 * __restrict__ W, int t )
{
    constexpr int R    = S < P4<N>::N4 ? 4 : 2;
    constexpr int Ns   = ipow4c( S );
    constexpr int M    = 4 / R;
    constexpr int T    = P4<N>::T;
    constexpr bool LST = ( S + 1 == P4<N>::NST );
    double2 out[4];
#pragma unroll
    for( int m = 0; m < M; ++m )
    {
        const int j  = t + T * m;
        const int jm = j & ( Ns - 1 );
        double2 u[R];
#pragma unroll
        for( int r = 0; r < R; ++r )
            u[r] = v[m + M * r];
        if constexpr( S > 0 )
        {
#pragma unroll
            for( int r = 1; r < R; ++r )
            {
                double2 w = W[( r * jm * ( N / ( Ns * R ) ) ) & ( N - 1 )];
                if( INV )
                    w.y = -w.y;
                u[r] = cmul( u[r], w );
            }
        }
        dft_small<R, INV>( u, nullptr, N );
        if constexpr( LST )
        {
            // Ns = N/R: output j + r*N/R = t + T*(m + M*r)
#pragma unroll
            for( int r = 0; r < R; ++r )
                out[m + M * r] = u[r];
        }
        else
        {
            const int idxD = ( j / Ns ) * Ns * R + jm;
#pragma unroll
            for( int r = 0; r < R; ++r )
                L[pad8( idxD + r * Ns )] = u[r];
        }
    }
    if constexpr( LST )
    {
#pragma unroll
        for( int k = 0; k < 4; ++k )
            v[k] = out[k];
    }
    else
    {
        __syncthreads();
#pragma unroll
        for( int k = 0; k < 4; ++k )
            v[k] = L[pad8( t + T * k )];
        __syncthreads();
        p4_stage<N, INV, S + 1>( v, L, W, t );
    }
}

template<int N, bool INV>
__device__ __forceinline__ void p4_fft_reg( double2 ( &v )[4], double2 * __restrict__ L,
                                            const double2 * __restrict__ W, int t )
{
    p4_stage<N, INV, 0>( v, L, W, t );
}

// Transform ending in LDS (natural order, L[pad8(i)]).
template<int N, bool INV>
__device__ __forceinline__ void p2_fft( double2 ( &v )[8], double2 * __restrict__ L, const double2 * __restrict__ tw,
                                        int t )
{
    p2_stage<N, INV, 0, false>( v, L, tw, t );
}

// Transform ending in registers: v[k] = X[t + T*k].  Every thread of the block must call it
// (it synchronises the block between stages); L may be reused right after it returns.
template<int N, bool INV>
__device__ __forceinline__ void p2_fft_reg( double2 ( &v )[8], double2 * __restrict__ L,
                                            const double2 * __restrict__ tw, int t )
{
    p2_stage<N, INV, 0, true>( v, L, tw, t );
}

} // namespace spirit_hip
