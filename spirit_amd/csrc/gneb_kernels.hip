// GNEB chain kernels (FP64), hand-written for gfx950.
//
// Restated reference code (paths relative to the reference checkout):
//   Method_GNEB::Calculate_Force / _Virtual   core/src/engine/Method_GNEB.cpp:88-391
//   Method_GNEB::Hook_Post_Iteration           core/src/engine/Method_GNEB.cpp:411-456
//   Manifoldmath (tangents, projections)       core/src/engine/Manifoldmath.cpp:23-209
//   Cubic_Hermite_Spline::Interpolate          core/src/utility/Cubic_Hermite_Spline.cpp:10-47
//   Method_Solver<VP>::Iteration               core/include/engine/Solver_VP.hpp:29-114
//   Solver iterations on a chain               core/include/engine/Solver_{Heun,Depondt,SIB,RK4}.hpp
//
// Chain-wide dot products are per-block partial sums reduced per image in a fixed order
// (deterministic; the order differs from the reference's OpenMP reductions, so results
// agree to rounding).  A dot with a normalised tangent is formed as (1/|tau|) * (F . tau_raw).
#include "gneb.hpp"
#include "llg_device.hpp"

namespace spirit_hip
{

// Sum of NV per-thread values over the block; thread 0 writes them to out[slot[k]].
// A block's partial sums are read by the last block of the same launch, which may run on
// another XCD (the XCDs' L2 caches are not coherent with each other): they are stored with
// device-scope stores, which write through to memory, so the arrival in pass_tail needs no
// device-wide cache write-back.
__device__ __forceinline__ void partial_store( double * p, double v )
{
    __hip_atomic_store( p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT );
}
// ... and read back with device-scope loads, which bypass the reading XCD's cache, so no
// acquire fence (an invalidation of that XCD's whole L2) is needed before them
__device__ __forceinline__ double partial_load( const double * p )
{
    return __hip_atomic_load( p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT );
}

template<int NV>
__device__ __forceinline__ void block_sums( double ( &v )[NV], double * out, const int ( &slot )[NV] )
{
    __shared__ double sh[NV][GNEB_THREADS / 64];
#pragma unroll
    for( int k = 0; k < NV; ++k )
        for( int off = 32; off > 0; off >>= 1 )
            v[k] += __shfl_down( v[k], off, 64 );
    const int w = threadIdx.x / 64, lane = threadIdx.x % 64;
    if( lane == 0 )
#pragma unroll
        for( int k = 0; k < NV; ++k )
            sh[k][w] = v[k];
    __syncthreads();
    if( threadIdx.x == 0 )
#pragma unroll
        for( int k = 0; k < NV; ++k )
        {
            double s = sh[k][0];
            for( int j = 1; j < GNEB_THREADS / 64; ++j )
                s += sh[k][j];
            partial_store( out + slot[k], s );
        }
}

__device__ __forceinline__ void block_max_store( double v, double * out )
{
    __shared__ double sh[GNEB_THREADS / 64];
    for( int off = 32; off > 0; off >>= 1 )
        v = fmax( v, __shfl_down( v, off, 64 ) );
    const int w = threadIdx.x / 64, lane = threadIdx.x % 64;
    if( lane == 0 )
        sh[w] = v;
    __syncthreads();
    if( threadIdx.x == 0 )
    {
        double m = sh[0];
        for( int j = 1; j < GNEB_THREADS / 64; ++j )
            m = fmax( m, sh[j] );
        partial_store( out, m );
    }
}

// Vectormath::angle, Vectormath.cpp:434-441
__device__ __forceinline__ double angle3( const double * a, const double * b )
{
    double r = dot3( a, b );
    r        = fmax( -1.0, fmin( 1.0, r ) );
    return acos( r );
}

// Rotation matrix about a unit axis (Rodrigues), as Eigen::AngleAxis::toRotationMatrix
__device__ __forceinline__ void rot_matrix( const double * axis, double angle, double R[3][3] )
{
    double s, c;
    sincos( angle, &s, &c );
    const double sa[3] = { s * axis[0], s * axis[1], s * axis[2] };
    const double ca[3] = { ( 1 - c ) * axis[0], ( 1 - c ) * axis[1], ( 1 - c ) * axis[2] };
    double t;
    t       = ca[0] * axis[1];
    R[0][1] = t - sa[2];
    R[1][0] = t + sa[2];
    t       = ca[0] * axis[2];
    R[0][2] = t + sa[1];
    R[2][0] = t - sa[1];
    t       = ca[1] * axis[2];
    R[1][2] = t - sa[0];
    R[2][1] = t + sa[0];
    for( int q = 0; q < 3; ++q )
        R[q][q] = ca[q] * axis[q] + c;
}

// normalised tangent of image img at spin i
__device__ __forceinline__ void tau_n( const GnebArgs & A, long long base, int i, const double * ch, double * t )
{
    load3( A.tau + base, A.nos, i, t );
    const double sc = ch[GC_SC];
#pragma unroll
    for( int q = 0; q < 3; ++q )
        t[q] *= sc;
}

// F_gradient projected orthogonal to the tangent (project_orthogonal, Manifoldmath.cpp:32-39)
__device__ __forceinline__ void f_orth( const GnebArgs & A, long long base, int i, const double * ch, double * fo,
                                        double * tn )
{
    load3( A.Fg + base, A.nos, i, fo );
    tau_n( A, base, i, ch, tn );
    const double x = ch[GC_PROJ];
#pragma unroll
    for( int q = 0; q < 3; ++q )
        fo[q] -= x * tn[q];
}

// path shortening vectors (Method_GNEB.cpp:200-229): normalised secants of the chain spins
__device__ __forceinline__ void shrink_secants( const GnebArgs & A, long long base, int i, const double * ch,
                                                double * tp, double * tm )
{
    const long long n3 = 3LL * A.nos;
    double xp[3], x0[3], xm[3];
    load3( A.xc + base + n3, A.nos, i, xp );
    load3( A.xc + base, A.nos, i, x0 );
    load3( A.xc + base - n3, A.nos, i, xm );
#pragma unroll
    for( int q = 0; q < 3; ++q )
    {
        tp[q] = 1 * xp[q];
        tp[q] += -1 * x0[q];
        tm[q] = 1 * x0[q];
        tm[q] += -1 * xm[q];
    }
}

// f_shrink after the two orthogonal projections (before normalisation)
__device__ __forceinline__ void shrink_vector( const GnebArgs & A, long long base, int i, const double * ch, int upto,
                                               double * f, double * tp2, double * tn )
{
    double tp[3], tm[3], fo[3];
    shrink_secants( A, base, i, ch, tp, tm );
    f_orth( A, base, i, ch, fo, tn );
    const double ntp = ch[GC_NTP], ntm = ch[GC_NTM];
#pragma unroll
    for( int q = 0; q < 3; ++q )
    {
        f[q] = 1 * ( tp[q] * ntp );
        f[q] += -1 * ( tm[q] * ntm );
    }
    const double ig = 1.0 / ch[GC_GNORM];
#pragma unroll
    for( int q = 0; q < 3; ++q )
        tp2[q] = fo[q] * ig;
    if( upto >= 1 )
    {
        const double x1 = ch[GC_X1];
#pragma unroll
        for( int q = 0; q < 3; ++q )
            f[q] -= x1 * tp2[q];
    }
    if( upto >= 2 )
    {
        const double x2 = ch[GC_X2];
#pragma unroll
        for( int q = 0; q < 3; ++q )
            f[q] -= x2 * tn[q];
    }
}

// F_total of image img at spin i (Method_GNEB.cpp:186-262 and 265-355)
__device__ __forceinline__ void total_force( const GnebArgs & A, int img, long long base, int i, const double * ch,
                                             double * F )
{
    const int noi = A.noi;
    F[0] = F[1] = F[2] = 0;
    if( img == 0 || img == noi - 1 )
    {
        if( !A.P.moving_endpoints )
            return;
        double fg[3], tn[3], ftr[3] = { 0, 0, 0 };
        load3( A.Fg + base, A.nos, i, fg );
        tau_n( A, base, i, ch, tn );
        if( A.P.translating_endpoints )
        {
            const double pt = ch[GC_PT];
#pragma unroll
            for( int q = 0; q < 3; ++q )
                ftr[q] = pt * tn[q];
        }
        const double rot = ch[GC_ROT], proj = ch[GC_PROJ], tc = ch[GC_TC];
#pragma unroll
        for( int q = 0; q < 3; ++q )
            F[q] = rot * ( fg[q] - proj * tn[q] ) + tc * tn[q] + ftr[q];
        return;
    }
    const int type = A.types[img];
    if( type == 1 ) // climbing: invert_parallel, Manifoldmath.cpp:41-47
    {
        double fg[3], tn[3];
        load3( A.Fg + base, A.nos, i, fg );
        tau_n( A, base, i, ch, tn );
        const double x2 = 2 * ch[GC_PROJ];
#pragma unroll
        for( int q = 0; q < 3; ++q )
            F[q] = fg[q] - x2 * tn[q];
    }
    else if( type == 2 ) // falling
        load3( A.Fg + base, A.nos, i, F );
    else if( type == 0 ) // normal
    {
        double fo[3], tn[3];
        f_orth( A, base, i, ch, fo, tn );
        const double d = ch[GC_D];
#pragma unroll
        for( int q = 0; q < 3; ++q )
        {
            F[q] = 1 * fo[q];
            F[q] += 1 * ( d * tn[q] );
        }
        if( A.P.path_shortening_constant > 0 )
        {
            double f[3], tp2[3], tn2[3];
            shrink_vector( A, base, i, ch, 2, f, tp2, tn2 );
            const double sc2 = ch[GC_SC2], scale = ch[GC_SCALE];
#pragma unroll
            for( int q = 0; q < 3; ++q )
                F[q] += 1 * ( ( f[q] * sc2 ) * scale );
        }
    }
}

template<int PHASE>
__device__ __forceinline__ void pass_tail( const GnebArgs & A );

// One per-spin pass over every image of the chain; with PHASE >= 0 the last block to finish
// runs that chain phase on the pass's partial sums (pass_tail), so a pass and the chain step
// that consumes it are one launch.
template<int PASS, int PHASE>
__global__ void __launch_bounds__( GNEB_THREADS ) k_gneb_pass( GnebArgs A )
{
    if( *A.stop )
        return;
    const int blk         = blockIdx.x;
    const int img         = blk / A.bpi;
    const int i           = ( blk - img * A.bpi ) * GNEB_THREADS + threadIdx.x;
    const bool act        = i < A.nos;
    const int nos         = A.nos;
    const long long base  = 3LL * img * nos;
    const long long n3    = 3LL * nos;
    const double * ch     = A.ch + img * GNEB_NCH;
    double * part         = A.part + (long long)blk * GNEB_NPART;
    const int noi         = A.noi;
    const bool interior   = img > 0 && img < noi - 1;

    if constexpr( PASS == GPASS_GRAD )
    {
        double v[2] = { 0, 0 };
        if( act )
        {
            const ImageDev & I = A.imgs[img];
            double g[3], e;
            local_gradient( I.ham, A.x + base, nos, i, ( A.h && I.ham.has_ddi ) ? A.h + base : nullptr, g, e );
            v[0] = e;
            double ef[3] = { -1 * g[0], -1 * g[1], -1 * g[2] };
            store3( A.eff + base, nos, i, ef );
            if( img > 0 )
            {
                double a[3], b[3];
                load3( A.x + base, nos, i, a );
                load3( A.x + base - n3, nos, i, b );
                const double t = angle3( a, b );
                v[1]           = t * t;
            }
        }
        const int slot[2] = { GP_E, GP_DIST };
        block_sums<2>( v, part, slot );
    }
    else if constexpr( PASS == GPASS_TANGENT )
    {
        double v[3] = { 0, 0, 0 };
        if( act )
        {
            double x[3], t[3], ef[3];
            load3( A.x + base, nos, i, x );
            load3( A.eff + base, nos, i, ef );
            if( interior )
            {
                double xp[3], xm[3], tp[3], tm[3];
                load3( A.x + base + n3, nos, i, xp );
                load3( A.x + base - n3, nos, i, xm );
#pragma unroll
                for( int q = 0; q < 3; ++q )
                {
                    tp[q] = 1 * xp[q];
                    tp[q] += -1 * x[q];
                    tm[q] = 1 * x[q];
                    tm[q] += -1 * xm[q];
                }
                const int mode = (int)ch[GC_MODE];
                if( mode == 1 )
                {
#pragma unroll
                    for( int q = 0; q < 3; ++q )
                        t[q] = 1 * tp[q];
                }
                else if( mode == 2 )
                {
#pragma unroll
                    for( int q = 0; q < 3; ++q )
                        t[q] = tm[q];
                }
                else
                {
                    const double cp = ch[GC_CP], cm = ch[GC_CM];
#pragma unroll
                    for( int q = 0; q < 3; ++q )
                    {
                        t[q] = cp * tp[q];
                        t[q] += cm * tm[q];
                    }
                }
                const double d = dot3( t, x );
#pragma unroll
                for( int q = 0; q < 3; ++q )
                    t[q] -= d * x[q];
            }
            else
            {
                // Geodesic_Tangent, Manifoldmath.cpp:78-108
                double mn[3], pl[3];
                if( img == 0 )
                {
                    load3( A.x + base, nos, i, mn );
                    load3( A.x + base + n3, nos, i, pl );
                }
                else
                {
                    load3( A.x + base - n3, nos, i, mn );
                    load3( A.x + base, nos, i, pl );
                }
                double axis[3];
                cross3( pl, mn, axis );
                const double eps = 1e-15;
                if( fabs( dot3( mn, pl ) + 1 ) < eps )
                {
                    const bool use_ex = fabs( x[0] - 1 ) > eps; // mid . ex
                    axis[0]           = use_ex ? 1.0 : 0.0;
                    axis[1]           = use_ex ? 0.0 : 1.0;
                    axis[2]           = 0.0;
                }
                cross3( x, axis, t );
            }
            store3( A.tau + base, nos, i, t );
            double fg[3] = { ef[0], ef[1], ef[2] };
            if( interior || A.P.moving_endpoints )
            {
                const double d = dot3( fg, x );
#pragma unroll
                for( int q = 0; q < 3; ++q )
                    fg[q] -= d * x[q];
            }
            store3( A.Fg + base, nos, i, fg );
            v[0] = dot3( t, t );
            v[1] = dot3( fg, t );
            v[2] = dot3( ef, t );
        }
        const int slot[3] = { GP_TT, GP_FT, GP_ET };
        block_sums<3>( v, part, slot );
    }
    else if constexpr( PASS == GPASS_TRANSLATE )
    {
        // launched over the blocks of image 0; one thread handles spin i of both endpoints
        double v[2] = { 0, 0 };
        if( act )
        {
            const long long bl = 0, br = 3LL * ( noi - 1 ) * nos;
            double sl[3], sr[3], gl[3], gr[3];
            load3( A.xc + bl, nos, i, sl );
            load3( A.xc + br, nos, i, sr );
            load3( A.Fg + bl, nos, i, gl );
            load3( A.Fg + br, nos, i, gr );
            double axis[3];
            cross3( sl, sr, axis );
            const double c     = dot3( sl, sr );
            const double angle = acos( c );
            normalize3( axis );
            double R[3][3];
            rot_matrix( axis, angle, R );
            if( fabs( c ) >= 1.0 )
                for( int a = 0; a < 3; ++a )
                    for( int b = 0; b < 3; ++b )
                        R[a][b] = a == b ? 1.0 : 0.0;
            double fl[3], fr[3];
#pragma unroll
            for( int a = 0; a < 3; ++a )
            {
                const double rr = R[a][0] * gr[0] + R[a][1] * gr[1] + R[a][2] * gr[2];
                const double rl = R[0][a] * gl[0] + R[1][a] * gl[1] + R[2][a] * gl[2];
                fl[a]           = -0.5 * ( gl[a] + rr );
                fr[a]           = -0.5 * ( rl + gr[a] );
            }
            store3( A.ftr, nos, i, fl );
            store3( A.ftr + n3, nos, i, fr );
            double tl[3], trr[3];
            tau_n( A, bl, i, A.ch, tl );
            tau_n( A, br, i, A.ch + ( noi - 1 ) * GNEB_NCH, trr );
            v[0] = dot3( fl, tl );
            v[1] = dot3( fr, trr );
        }
        const int slot[2] = { GP_A, GP_B };
        block_sums<2>( v, part, slot );
    }
    else if constexpr( PASS == GPASS_SHRINK1 || PASS == GPASS_SHRINK2 || PASS == GPASS_SHRINK3
                       || PASS == GPASS_SHRINK4 )
    {
        double v[3]        = { 0, 0, 0 };
        const bool normal  = interior && A.types[img] == 0;
        if( act && normal )
        {
            if constexpr( PASS == GPASS_SHRINK1 )
            {
                double tp[3], tm[3], fo[3], tn[3];
                shrink_secants( A, base, i, ch, tp, tm );
                f_orth( A, base, i, ch, fo, tn );
                v[0] = dot3( tp, tp );
                v[1] = dot3( tm, tm );
                v[2] = dot3( fo, fo );
            }
            else
            {
                double f[3], tp2[3], tn[3];
                shrink_vector( A, base, i, ch, PASS - GPASS_SHRINK2, f, tp2, tn );
                if constexpr( PASS == GPASS_SHRINK2 )
                    v[0] = dot3( f, tp2 );
                else if constexpr( PASS == GPASS_SHRINK3 )
                    v[0] = dot3( f, tn );
                else
                    v[0] = dot3( f, f );
            }
        }
        const int slot[3] = { GP_A, GP_B, GP_C };
        block_sums<3>( v, part, slot );
    }
    else if constexpr( PASS == GPASS_FORCE_VP || PASS == GPASS_FORCE_FV )
    {
        double v[2] = { 0, 0 };
        if( act )
        {
            double F[3];
            total_force( A, img, base, i, ch, F );
            store3( A.Ft + base, nos, i, F );
            if constexpr( PASS == GPASS_FORCE_VP )
            {
                // v += 0.5/m (f_prev + f), Solver_VP.hpp:63-71
                double vel[3], fp[3];
                load3( A.v + base, nos, i, vel );
                load3( A.fprev + base, nos, i, fp );
                const double c = 0.5 / A.P.m;
#pragma unroll
                for( int q = 0; q < 3; ++q )
                    vel[q] += c * ( fp[q] + F[q] );
                store3( A.v + base, nos, i, vel );
                v[0] = dot3( vel, F );
                v[1] = dot3( F, F );
            }
            else if( interior || A.P.moving_endpoints )
            {
                // Calculate_Force_Virtual, Method_GNEB.cpp:360-391
                const double dtg = A.imgs[img].llg.dt * gamma_e / mu_B;
                double x[3], c[3];
                load3( A.x + base, nos, i, x );
                cross3( x, F, c );
#pragma unroll
                for( int q = 0; q < 3; ++q )
                    c[q] *= dtg;
                store3( A.fv + base, nos, i, c );
            }
        }
        if constexpr( PASS == GPASS_FORCE_VP )
        {
            const int slot[2] = { GP_A, GP_B };
            block_sums<2>( v, part, slot );
        }
    }
    else if constexpr( PASS == GPASS_VP_UPDATE || PASS == GPASS_HOOK )
    {
        double t2 = 0;
        if( act )
        {
            double F[3], x[3];
            load3( A.Ft + base, nos, i, F );
            double * xw = const_cast<double *>( A.x );
            load3( xw + base, nos, i, x );
            if constexpr( PASS == GPASS_VP_UPDATE )
            {
                // Solver_VP.hpp:86-111
                const double * c0  = A.ch;
                const double dt    = A.imgs[img].llg.dt;
                const double ratio = c0[GC_RATIO];
                double vel[3];
                if( c0[GC_VZERO] != 0 )
                    vel[0] = vel[1] = vel[2] = 0;
                else
#pragma unroll
                    for( int q = 0; q < 3; ++q )
                        vel[q] = F[q] * ratio;
                store3( A.v + base, nos, i, vel );
                const double cf = 0.5 / A.P.m * dt;
#pragma unroll
                for( int q = 0; q < 3; ++q )
                    x[q] = x[q] + dt * vel[q] + cf * F[q];
                normalize3( x );
                store3( xw + base, nos, i, x );
            }
            // Hook_Post_Iteration: MaxTorque_on_Image projects F_total, then the forces are projected
            const double d = dot3( F, x );
#pragma unroll
            for( int q = 0; q < 3; ++q )
                F[q] -= d * x[q];
            if constexpr( PASS == GPASS_VP_UPDATE )
                store3( A.fprev + base, nos, i, F );
            t2 = dot3( F, F );
        }
        block_max_store( t2, part + GP_C );
    }
    if constexpr( PHASE >= 0 )
        pass_tail<PHASE>( A );
}

// Cubic Hermite interpolation point k of the chain (Cubic_Hermite_Spline.cpp:10-47)
__device__ void hermite_point( const double * ch, int noi, int n_interp, int k, double & ox, double & op )
{
    const int npts = noi + ( noi - 1 ) * n_interp;
    if( k == npts - 1 )
    {
        ox = ch[( noi - 1 ) * GNEB_NCH + GC_RX];
        op = ch[( noi - 1 ) * GNEB_NCH + GC_E];
        return;
    }
    const int i = k / ( n_interp + 1 ), j = k % ( n_interp + 1 );
    const double * a = ch + i * GNEB_NCH;
    const double * b = ch + ( i + 1 ) * GNEB_NCH;
    const double x0 = a[GC_RX], x1 = b[GC_RX], p0 = a[GC_E], p1 = b[GC_E], m0 = a[GC_DEDRX], m1 = b[GC_DEDRX];
    const double t   = j / (double)( n_interp + 1 );
    const double h00 = 2 * pow( t, 3.0 ) - 3 * pow( t, 2.0 ) + 1;
    const double h10 = -2 * pow( t, 3.0 ) + 3 * pow( t, 2.0 );
    const double h01 = pow( t, 3.0 ) - 2 * pow( t, 2.0 ) + t;
    const double h11 = pow( t, 3.0 ) - pow( t, 2.0 );
    ox               = x0 + t * ( x1 - x0 );
    op               = h00 * p0 + h10 * p1 + h01 * m0 * ( x0 - x1 ) + h11 * m1 * ( x0 - x1 );
}

template<int PHASE>
__device__ __forceinline__ void chain_phase( const GnebArgs & A, double * ch, const double ( *red )[256] );

// Partial-sum slots a chain phase consumes and how they combine (sums, or maxima for the hook).
template<int PHASE>
struct PhaseSlots
{
    static constexpr bool MAXRED = PHASE == GPH_HOOK;
    static constexpr int NS      = ( PHASE == GPH_DIST || PHASE == GPH_TRANSLATE || PHASE == GPH_VP ) ? 2
                                   : ( PHASE == GPH_TANGENT || PHASE == GPH_SHRINK1 )            ? 3
                                                                                                 : 1;
    __device__ static int slot( int k )
    {
        if( PHASE == GPH_DIST )
            return k == 0 ? GP_E : GP_DIST;
        if( PHASE == GPH_TANGENT )
            return k == 0 ? GP_TT : k == 1 ? GP_FT : GP_ET;
        if( PHASE == GPH_TRANSLATE || PHASE == GPH_VP )
            return k == 0 ? GP_A : GP_B;
        if( PHASE == GPH_SHRINK1 )
            return k == 0 ? GP_A : k == 1 ? GP_B : GP_C;
        if( PHASE == GPH_HOOK )
            return GP_C;
        return GP_A;
    }
};

// The sums of one image's block partials (wave 0 of the image's last block): lane b adds the
// partials of blocks b, b+64, ... in block order, then a shuffle tree; stored to isum[img]
// with device-scope stores for the chain's last image.
template<int PHASE>
__device__ __forceinline__ void image_sums( const GnebArgs & A, int img )
{
    using PS       = PhaseSlots<PHASE>;
    const int lane = threadIdx.x;
    if( threadIdx.x >= 64 )
        return;
    double s[3] = { 0, 0, 0 };
    for( int b = lane; b < A.bpi; b += 64 )
    {
        const double * p = A.part + (long long)( img * A.bpi + b ) * GNEB_NPART;
#pragma unroll
        for( int k = 0; k < PS::NS; ++k )
        {
            const double x = partial_load( p + PS::slot( k ) );
            s[k]           = PS::MAXRED ? fmax( s[k], x ) : s[k] + x;
        }
    }
#pragma unroll
    for( int k = 0; k < PS::NS; ++k )
        for( int off = 32; off > 0; off >>= 1 )
        {
            const double o = __shfl_down( s[k], off, 64 );
            s[k]           = PS::MAXRED ? fmax( s[k], o ) : s[k] + o;
        }
    if( lane == 0 )
        for( int k = 0; k < PS::NS; ++k )
            partial_store( A.isum + img * GNEB_ISUM + k, s[k] );
}

// The chain step that follows a per-spin pass, run by the block that completes the last
// image (every thread of the block calls it): the images' sums, then the chain logic of the
// phase.  The chain scalars are staged in LDS (chains of up to GNEB_CH_LDS images): the
// sequential parts of the logic then pay LDS latency instead of L2 latency per dependent read.
template<int PHASE>
__device__ void chain_tail( const GnebArgs & A, int n_red_images )
{
    using PS = PhaseSlots<PHASE>;
    extern __shared__ double ch_lds[];
    const int noi       = A.noi;
    const bool staged   = noi <= GNEB_CH_LDS;
    const int n_scalars = noi * GNEB_NCH;
    if( staged )
        for( int e = threadIdx.x; e < n_scalars; e += blockDim.x )
            ch_lds[e] = A.ch[e];
    __shared__ double red[3][256];
    for( int img = threadIdx.x; img < n_red_images; img += blockDim.x )
#pragma unroll
        for( int k = 0; k < PS::NS; ++k )
            red[k][img] = partial_load( A.isum + img * GNEB_ISUM + k );
    __syncthreads();
    chain_phase<PHASE>( A, staged ? ch_lds : A.ch, red );
    __syncthreads();
    if( staged )
        for( int e = threadIdx.x; e < n_scalars; e += blockDim.x )
            A.ch[e] = ch_lds[e];
}

// Arrival of a block at the end of a pass, in two levels so that no counter takes more than
// one image's worth of atomics: thread 0 waits until its device-scope partial stores are
// performed and counts the block in its image; the image's last block sums the image's
// partials and counts the image in the chain; the chain's last image runs the chain phase.
// Counters are rearmed by the blocks that completed them.  A device-wide release fence in
// every block instead would write back each XCD's whole L2 once per block; the partials and
// image sums are read with device-scope loads (partial_load), so no acquire fence invalidates
// the reading block's L2 either.  Every thread of the block calls it, after the block's
// partial slots were stored by thread 0 (partial_store).
template<int PHASE>
__device__ __forceinline__ void pass_tail( const GnebArgs & A )
{
    __shared__ int last;
    const int img      = blockIdx.x / A.bpi;
    const int n_images = gridDim.x / A.bpi; // the translation pass runs over the blocks of image 0 only
    int * img_count    = A.count + GNEB_COUNT_STRIDE * ( 1 + img );
    if( threadIdx.x == 0 )
    {
        __builtin_amdgcn_fence( __ATOMIC_RELEASE, "workgroup" );
        __builtin_amdgcn_s_waitcnt( 0 ); // vmcnt = expcnt = lgkmcnt = 0: the partial stores are complete
        last = __hip_atomic_fetch_add( img_count, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT ) == A.bpi - 1 ? 1 : 0;
    }
    __syncthreads();
    if( !last )
        return;
    image_sums<PHASE>( A, img );
    __syncthreads();
    if( threadIdx.x == 0 )
    {
        __hip_atomic_store( img_count, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT );
        __builtin_amdgcn_fence( __ATOMIC_RELEASE, "workgroup" );
        __builtin_amdgcn_s_waitcnt( 0 ); // the image sums (stored by lane 0 of wave 0, this thread) are complete
        last = __hip_atomic_fetch_add( A.count, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT ) == n_images - 1 ? 1
                                                                                                                 : 0;
    }
    __syncthreads();
    if( !last )
        return;
    chain_tail<PHASE>( A, n_images );
    if( threadIdx.x == 0 )
        __hip_atomic_store( A.count, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT );
}

// The chain logic of one phase, run by the whole block: per-image work on one thread per
// image, the sequential parts (the reaction-coordinate prefix sum, chain-wide sums and
// maxima, the spline, the endpoints) on thread 0 in the reference's order.  Every thread
// calls it (it synchronises the block).
template<int PHASE>
__device__ __forceinline__ void chain_phase( const GnebArgs & A, double * ch, const double ( *red )[256] )
{
    const int noi           = A.noi;
    const int tid           = threadIdx.x, nth = blockDim.x;
    const GnebParamsDev & P = A.P;
    auto C = [&]( int img, int slot ) -> double & { return ch[img * GNEB_NCH + slot]; };

    if constexpr( PHASE == GPH_DIST )
    {
        // Method_GNEB.cpp:96-126 and Manifoldmath::Tangents weights, :144-199
        __shared__ double dist[256];
        for( int img = tid; img < noi; img += nth )
        {
            C( img, GC_E )     = red[0][img];
            C( img, GC_DIST2 ) = red[1][img];
            dist[img]          = sqrt( red[1][img] );
            if( img >= 1 && img < noi - 1 )
            {
                const double E_mid = red[0][img], E_plus = red[0][img + 1], E_minus = red[0][img - 1];
                double mode = 0, cp = 1, cm = 1;
                if( ( E_plus < E_mid && E_mid > E_minus ) || ( E_plus > E_mid && E_mid < E_minus ) )
                {
                    const double E_max = fmax( fabs( E_plus - E_mid ), fabs( E_minus - E_mid ) );
                    const double E_min = fmin( fabs( E_plus - E_mid ), fabs( E_minus - E_mid ) );
                    cp                 = E_plus > E_minus ? E_max : E_min;
                    cm                 = E_plus > E_minus ? E_min : E_max;
                }
                else if( E_plus > E_mid && E_mid > E_minus )
                    mode = 1;
                else if( E_plus < E_mid && E_mid < E_minus )
                    mode = 2;
                C( img, GC_MODE ) = mode;
                C( img, GC_CP )   = cp;
                C( img, GC_CM )   = cm;
            }
        }
        __syncthreads();
        if( tid == 0 )
        {
            C( 0, GC_RX ) = 0;
            for( int img = 1; img < noi; ++img )
            {
                C( img, GC_RX ) = C( img - 1, GC_RX ) + dist[img];
                if( C( img, GC_RX ) - C( img - 1, GC_RX ) < 1e-10 )
                {
                    *A.stop = 2;
                    break;
                }
            }
        }
    }
    else if constexpr( PHASE == GPH_TANGENT )
    {
        for( int img = tid; img < noi; img += nth )
        {
            const double sc    = 1.0 / sqrt( red[0][img] );
            C( img, GC_SC )    = sc;
            C( img, GC_PROJ )  = sc * red[1][img];
            C( img, GC_DEDRX ) = sc * red[2][img];
        }
        __syncthreads();
        // spring coefficients, Method_GNEB.cpp:131-167 and 232-236
        if( P.spring_force_ratio > 0 )
        {
            if( tid == 0 )
            {
                const double ratio_E = fmin( 1.0, P.spring_force_ratio ), ratio_Rx = 1 - ratio_E;
                const int n_interp = 20;
                const int npts     = noi + ( noi - 1 ) * n_interp;
                double xmin = 0, xmax = 0, pmin = 0, pmax = 0;
                for( int k = 0; k < npts; ++k )
                {
                    double ox, op;
                    hermite_point( ch, noi, n_interp, k, ox, op );
                    if( k == 0 || ox > xmax )
                        xmax = ox;
                    if( k == 0 || ox < xmin )
                        xmin = ox;
                    if( k == 0 || op > pmax )
                        pmax = op;
                    if( k == 0 || op < pmin )
                        pmin = op;
                }
                const double range_Rx = xmax - xmin, range_E = pmax - pmin;
                C( 0, GC_LEN ) = 0;
                for( int img = 1; img < noi; ++img )
                {
                    double len = 0;
                    for( int i = 1; i <= n_interp; ++i )
                    {
                        const int idx = ( img - 1 ) * ( n_interp + 1 ) + i;
                        double x0, p0, x1, p1;
                        hermite_point( ch, noi, n_interp, idx - 1, x0, p0 );
                        hermite_point( ch, noi, n_interp, idx, x1, p1 );
                        const double dRx = ratio_Rx * ( x1 - x0 ) / range_Rx;
                        const double dE  = ratio_E * ( p1 - p0 ) / range_E;
                        len += sqrt( dRx * dRx + dE * dE );
                    }
                    C( img, GC_LEN ) = len * range_Rx;
                }
                for( int img = 1; img < noi - 1; ++img )
                    C( img, GC_D ) = P.spring_constant * ( C( img + 1, GC_LEN ) - C( img, GC_LEN ) );
            }
        }
        else
            for( int img = tid + 1; img < noi - 1; img += nth )
                C( img, GC_D )
                    = P.spring_constant * ( C( img + 1, GC_RX ) - 2 * C( img, GC_RX ) + C( img - 1, GC_RX ) );
        // moving endpoints, Method_GNEB.cpp:265-355
        if( P.moving_endpoints && tid == 0 )
        {
            double rot = 1.0;
            if( P.escape_first && C( 0, GC_PROJ ) > C( noi - 1, GC_PROJ ) )
                rot = 0.0;
            for( int e = 0; e < 2; ++e )
            {
                const int img    = e == 0 ? 0 : noi - 1;
                const double d0  = e == 0 ? P.delta_Rx_left : P.delta_Rx_right;
                const double dRx = e == 0 ? C( 1, GC_RX ) - C( 0, GC_RX ) : C( noi - 1, GC_RX ) - C( noi - 2, GC_RX );
                const double k   = ( e == 0 ? 1.0 : -1.0 ) * P.spring_constant;
                C( img, GC_ROT ) = rot;
                C( img, GC_TC )  = k * ( dRx - d0 );
                C( img, GC_PT )  = 0;
            }
        }
    }
    else if constexpr( PHASE == GPH_TRANSLATE )
    {
        if( tid == 0 )
        {
            C( 0, GC_PT )       = red[0][0];
            C( noi - 1, GC_PT ) = red[1][0];
        }
    }
    else if constexpr( PHASE == GPH_SHRINK1 )
    {
        for( int img = tid + 1; img < noi - 1; img += nth )
        {
            C( img, GC_NTP )   = 1.0 / sqrt( red[0][img] );
            C( img, GC_NTM )   = 1.0 / sqrt( red[1][img] );
            C( img, GC_GNORM ) = sqrt( red[2][img] );
        }
    }
    else if constexpr( PHASE == GPH_SHRINK2 )
    {
        for( int img = tid + 1; img < noi - 1; img += nth )
            C( img, GC_X1 ) = red[0][img];
    }
    else if constexpr( PHASE == GPH_SHRINK3 )
    {
        for( int img = tid + 1; img < noi - 1; img += nth )
            C( img, GC_X2 ) = red[0][img];
    }
    else if constexpr( PHASE == GPH_SHRINK4 )
    {
        for( int img = tid + 1; img < noi - 1; img += nth )
        {
            C( img, GC_SC2 )   = 1.0 / sqrt( red[0][img] );
            C( img, GC_SCALE ) = fmax( C( img, GC_GNORM ), A.nos * P.path_shortening_constant );
        }
    }
    else if constexpr( PHASE == GPH_VP )
    {
        // Solver_VP.hpp:73-93 (chain-wide sums in image order)
        if( tid == 0 )
        {
            double proj = 0, fn2 = 0;
            for( int img = 0; img < noi; ++img )
            {
                proj += red[0][img];
                fn2 += red[1][img];
            }
            C( 0, GC_RATIO ) = proj / fn2;
            C( 0, GC_VZERO ) = proj <= 0 ? 1.0 : 0.0;
        }
    }
    else if constexpr( PHASE == GPH_HOOK )
    {
        // Hook_Post_Iteration, Method_GNEB.cpp:411-432; convergence, Method_GNEB.cpp:393-397
        for( int img = tid; img < noi; img += nth )
            C( img, GC_TORQUE ) = sqrt( red[0][img] );
        __syncthreads();
        if( tid == 0 )
        {
            double mt = 0;
            for( int img = 0; img < noi; ++img )
                if( C( img, GC_TORQUE ) > mt )
                    mt = C( img, GC_TORQUE );
            C( 0, GC_MAXT ) = mt;
            *A.iter += 1;
            if( P.check_convergence && mt < P.force_convergence )
                *A.stop = 1;
        }
    }
}

// Solver stages on the chain (virtual forces already computed), restating the image loops
// of Solver_Heun.hpp:30-83, Solver_Depondt.hpp:29-77, Solver_SIB.hpp:22-52, Solver_RK4.hpp:41-147.
template<int SOLVER, int STAGE>
__global__ void __launch_bounds__( GNEB_THREADS ) k_gneb_stage( GnebSolverArgs S )
{
    if( *S.stop )
        return;
    const int blk        = blockIdx.x;
    const int img        = blk / S.bpi;
    const int i          = ( blk - img * S.bpi ) * GNEB_THREADS + threadIdx.x;
    const int nos        = S.nos;
    const long long base = 3LL * img * nos;
    if( i >= nos )
        return;
    double x[3], f[3];
    load3( S.x + base, nos, i, x );
    load3( S.fv + base, nos, i, f );
    if constexpr( SOLVER == SOLVER_HEUN )
    {
        if constexpr( STAGE == 0 )
        {
            double t[3], p[3];
            neg_cross3( x, f, t );
            store3( S.tmp + base, nos, i, t );
#pragma unroll
            for( int q = 0; q < 3; ++q )
            {
                p[q] = 1 * x[q];
                p[q] += 1 * t[q];
            }
            normalize3( p );
            store3( S.pred + base, nos, i, p );
        }
        else
        {
            double t[3], p[3], t1[3];
            load3( S.tmp + base, nos, i, t );
            load3( S.pred + base, nos, i, p );
#pragma unroll
            for( int q = 0; q < 3; ++q )
            {
                t[q] *= 0.5;
                t[q] += 1 * x[q];
            }
            neg_cross3( p, f, t1 );
#pragma unroll
            for( int q = 0; q < 3; ++q )
                t[q] += 0.5 * t1[q];
            normalize3( t );
            store3( S.x + base, nos, i, t );
        }
    }
    else if constexpr( SOLVER == SOLVER_DEPONDT )
    {
        if constexpr( STAGE == 0 )
        {
            double axis[3] = { 1 * f[0], 1 * f[1], 1 * f[2] }, p[3];
            const double angle = sqrt( dot3( f, f ) );
            normalize3( axis );
            rotate3( x, axis, angle, p );
            store3( S.pred + base, nos, i, p );
        }
        else
        {
            double f0[3], h[3], o[3];
            load3( S.fv0 + base, nos, i, f0 );
#pragma unroll
            for( int q = 0; q < 3; ++q )
            {
                h[q] = 0.5 * f0[q];
                h[q] += 0.5 * f[q];
            }
            const double angle = sqrt( dot3( h, h ) );
            normalize3( h );
            rotate3( x, h, angle, o );
            store3( S.x + base, nos, i, o );
        }
    }
    else if constexpr( SOLVER == SOLVER_SIB )
    {
        double o[3];
        sib3( x, f, o );
        if constexpr( STAGE == 0 )
        {
#pragma unroll
            for( int q = 0; q < 3; ++q )
            {
                o[q] += 1 * x[q];
                o[q] *= 0.5;
            }
            store3( S.pred + base, nos, i, o );
        }
        else
            store3( S.x + base, nos, i, o );
    }
    else if constexpr( SOLVER == SOLVER_RK4 )
    {
        double c[3], k[3], acc[3];
        if( STAGE == 0 )
            c[0] = x[0], c[1] = x[1], c[2] = x[2];
        else
            load3( S.pred + base, nos, i, c );
        neg_cross3( c, f, k );
        const double wk = ( STAGE == 0 || STAGE == 3 ) ? 1.0 / 6.0 : 1.0 / 3.0;
        if( STAGE == 0 )
#pragma unroll
            for( int q = 0; q < 3; ++q )
                acc[q] = 1 * x[q];
        else
            load3( S.tmp + base, nos, i, acc );
#pragma unroll
        for( int q = 0; q < 3; ++q )
            acc[q] += wk * k[q];
        if( STAGE < 3 )
        {
            store3( S.tmp + base, nos, i, acc );
            const double cp = STAGE == 2 ? 1.0 : 0.5;
            double p[3];
#pragma unroll
            for( int q = 0; q < 3; ++q )
            {
                p[q] = 1 * x[q];
                p[q] += cp * k[q];
            }
            normalize3( p );
            store3( S.pred + base, nos, i, p );
        }
        else
        {
            normalize3( acc );
            store3( S.x + base, nos, i, acc );
        }
    }
}

// ---------------------------------------------------------------------------
// host launchers
// ---------------------------------------------------------------------------
template<int PASS, int PHASE>
static void launch_pass( const GnebArgs & A, dim3 grid, hipStream_t st )
{
    const size_t lds = ( PHASE >= 0 && A.noi <= GNEB_CH_LDS ) ? sizeof( double ) * A.noi * GNEB_NCH : 0;
    hipLaunchKernelGGL( ( k_gneb_pass<PASS, PHASE> ), grid, dim3( GNEB_THREADS ), lds, st, A );
}

void gneb_launch_pass( int pass, const GnebArgs & A, hipStream_t st )
{
    const dim3 all( A.n_blocks ), first( A.bpi );
    switch( pass )
    {
        case GPASS_GRAD: launch_pass<GPASS_GRAD, GPH_DIST>( A, all, st ); break;
        case GPASS_TANGENT: launch_pass<GPASS_TANGENT, GPH_TANGENT>( A, all, st ); break;
        case GPASS_TRANSLATE: launch_pass<GPASS_TRANSLATE, GPH_TRANSLATE>( A, first, st ); break;
        case GPASS_SHRINK1: launch_pass<GPASS_SHRINK1, GPH_SHRINK1>( A, all, st ); break;
        case GPASS_SHRINK2: launch_pass<GPASS_SHRINK2, GPH_SHRINK2>( A, all, st ); break;
        case GPASS_SHRINK3: launch_pass<GPASS_SHRINK3, GPH_SHRINK3>( A, all, st ); break;
        case GPASS_SHRINK4: launch_pass<GPASS_SHRINK4, GPH_SHRINK4>( A, all, st ); break;
        case GPASS_FORCE_VP: launch_pass<GPASS_FORCE_VP, GPH_VP>( A, all, st ); break;
        case GPASS_FORCE_FV: launch_pass<GPASS_FORCE_FV, -1>( A, all, st ); break;
        case GPASS_VP_UPDATE: launch_pass<GPASS_VP_UPDATE, GPH_HOOK>( A, all, st ); break;
        case GPASS_HOOK: launch_pass<GPASS_HOOK, GPH_HOOK>( A, all, st ); break;
        default: break;
    }
}

template<int SOLVER, int STAGE>
static void launch_gs( const GnebSolverArgs & S, hipStream_t st )
{
    hipLaunchKernelGGL( ( k_gneb_stage<SOLVER, STAGE> ), dim3( S.noi * S.bpi ), dim3( GNEB_THREADS ), 0, st, S );
}

void gneb_launch_solver_stage( int solver, int stage, const GnebSolverArgs & S, hipStream_t st )
{
    switch( solver )
    {
        case SOLVER_HEUN: stage == 0 ? launch_gs<SOLVER_HEUN, 0>( S, st ) : launch_gs<SOLVER_HEUN, 1>( S, st ); break;
        case SOLVER_DEPONDT:
            stage == 0 ? launch_gs<SOLVER_DEPONDT, 0>( S, st ) : launch_gs<SOLVER_DEPONDT, 1>( S, st );
            break;
        case SOLVER_SIB: stage == 0 ? launch_gs<SOLVER_SIB, 0>( S, st ) : launch_gs<SOLVER_SIB, 1>( S, st ); break;
        case SOLVER_RK4:
            if( stage == 0 )
                launch_gs<SOLVER_RK4, 0>( S, st );
            else if( stage == 1 )
                launch_gs<SOLVER_RK4, 1>( S, st );
            else if( stage == 2 )
                launch_gs<SOLVER_RK4, 2>( S, st );
            else
                launch_gs<SOLVER_RK4, 3>( S, st );
            break;
        default: break;
    }
}

} // namespace spirit_hip
