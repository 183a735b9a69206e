// Discrete adjoint and tangent-linear kernels of the Depondt and Heun steps (FP64), for gfx950.
//
// The forward maps restated here are those of the fused stage kernels (llg_kernels.hip),
// themselves restating core/include/engine/Solver_Depondt.hpp:29-77,
// core/include/engine/Solver_Heun.hpp:30-83 and core/src/engine/Method_LLG.cpp:132-226.
// Derivatives are of the mathematical maps (FP rounding of the forward is not differentiated).
//
// Rotation by a vector h (the Depondt update): rot(h, v) = c v + a (h x v) + b (h.v) h with
// theta = |h|, c = cos theta, a = sin theta / theta, b = (1 - cos theta) / theta^2 -- the
// Rodrigues formula with axis h/|h| and angle |h|, written so that it is smooth at h = 0.
// With ap = a'(theta)/theta and bp = b'(theta)/theta:
//   d rot / dv     = R(h),  R(h)^T = R(-h)
//   (d rot / dh)^T mu = [ -a (v.mu) + ap ((h x v).mu) + bp (h.v)(h.mu) ] h
//                       + a (v x mu) + b ( (h.v) mu + (h.mu) v )
#include "adjoint.hpp"
#include "llg_device.hpp"

namespace spirit_hip
{

struct RotC
{
    double c, a, b, ap, bp;
};

__device__ __forceinline__ RotC rot_coeffs( double th )
{
    RotC r;
    if( th < 1e-2 )
    {
        // Taylor series (relative truncation error < 1e-16 for theta < 1e-2)
        const double t2 = th * th;
        r.c             = cos( th );
        r.a             = 1 - t2 / 6 * ( 1 - t2 / 20 * ( 1 - t2 / 42 ) );
        r.b             = 0.5 * ( 1 - t2 / 12 * ( 1 - t2 / 30 * ( 1 - t2 / 56 ) ) );
        r.ap            = -1.0 / 3 + t2 / 30 - t2 * t2 / 840;
        r.bp            = -1.0 / 12 + t2 / 180 - t2 * t2 / 6720;
    }
    else
    {
        double s, c;
        sincos( th, &s, &c );
        r.c  = c;
        r.a  = s / th;
        r.b  = ( 1 - c ) / ( th * th );
        r.ap = ( th * c - s ) / ( th * th * th );
        r.bp = ( th * s - 2 * ( 1 - c ) ) / ( th * th * th * th );
    }
    return r;
}

// R(h)^T mu = R(-h) mu
__device__ __forceinline__ void rot_T_v( const double * h, const RotC & r, const double * mu, double * out )
{
    double hx[3];
    cross3( h, mu, hx );
    const double hm = dot3( h, mu );
#pragma unroll
    for( int q = 0; q < 3; ++q )
        out[q] = r.c * mu[q] - r.a * hx[q] + r.b * hm * h[q];
}

// R(h) v
__device__ __forceinline__ void rot_v( const double * h, const RotC & r, const double * v, double * out )
{
    double hx[3];
    cross3( h, v, hx );
    const double hv = dot3( h, v );
#pragma unroll
    for( int q = 0; q < 3; ++q )
        out[q] = r.c * v[q] + r.a * hx[q] + r.b * hv * h[q];
}

// (d rot(h, v) / dh)^T mu
__device__ __forceinline__ void rot_T_h( const double * h, const RotC & r, const double * v, const double * mu,
                                         double * out )
{
    double hxv[3], vxm[3];
    cross3( h, v, hxv );
    cross3( v, mu, vxm );
    const double hv = dot3( h, v ), hm = dot3( h, mu );
    const double w  = -r.a * dot3( v, mu ) + r.ap * dot3( hxv, mu ) + r.bp * hv * hm;
#pragma unroll
    for( int q = 0; q < 3; ++q )
        out[q] = w * h[q] + r.a * vxm[q] + r.b * ( hv * mu[q] + hm * v[q] );
}

// (d rot(h, v) / dh) dh
__device__ __forceinline__ void rot_d_h( const double * h, const RotC & r, const double * v, const double * dh,
                                         double * out )
{
    double hxv[3], dhxv[3];
    cross3( h, v, hxv );
    cross3( dh, v, dhxv );
    const double hdh = dot3( h, dh ), hv = dot3( h, v ), dhv = dot3( dh, v );
#pragma unroll
    for( int q = 0; q < 3; ++q )
        out[q] = -r.a * hdh * v[q] + r.ap * hdh * hxv[q] + r.a * dhxv[q] + r.bp * hdh * hv * h[q]
                 + r.b * ( dhv * h[q] + hv * dh[q] );
}

// dtg / mu of the dynamic virtual force (Method_LLG.cpp:177-179)
__device__ __forceinline__ double vf_scale( const LLGDev & L, double mu )
{
    const double a = L.damping;
    return L.dt * gamma_e / mu_B / ( 1 + a * a ) / mu;
}

// adjoint of fv = k (F + alpha s x F) + xi + alpha s x xi for the input adjoint m = dL/dfv:
// dL/dF = k (m - alpha s x m);  dL/ds += k alpha (F x m) + alpha (xi x m)
__device__ __forceinline__ void vf_adjoint( double k, double alpha, const double * s, const double * F,
                                            const double * xi, const double * m, double * mF, double * ms )
{
    double sxm[3], Fxm[3], xxm[3];
    cross3( s, m, sxm );
    cross3( F, m, Fxm );
    cross3( xi, m, xxm );
#pragma unroll
    for( int q = 0; q < 3; ++q )
    {
        mF[q] = k * ( m[q] - alpha * sxm[q] );
        ms[q] += k * alpha * Fxm[q] + alpha * xxm[q];
    }
}

__device__ __forceinline__ void thermal_xi( const LLGDev & L, double mu, uint64_t step, int i, double * xi )
{
    xi[0] = xi[1] = xi[2] = 0;
    if( L.thermal && !L.direct_minimization )
        thermal_field( L, mu, step, i, nullptr, xi );
}

// normalisation n = u/|u|: dn = (du - (n.du) n)/|u|, and the same projection for the adjoint
__device__ __forceinline__ void norm_d( const double * u, const double * du, double * out )
{
    const double nu = sqrt( dot3( u, u ) );
    double n[3]     = { u[0] / nu, u[1] / nu, u[2] / nu };
    const double d  = dot3( n, du );
#pragma unroll
    for( int q = 0; q < 3; ++q )
        out[q] = ( du[q] - d * n[q] ) / nu;
}

// ---------------------------------------------------------------------------
// backward kernels
// ---------------------------------------------------------------------------

// predictor s' of the step, recomputed exactly as the forward stage kernel does
template<int SOLVER>
__global__ void __launch_bounds__( ADJ_THREADS ) k_adj_predictor( AdjArgs A )
{
    const int i = blockIdx.x * ADJ_THREADS + threadIdx.x;
    if( i >= A.nos )
        return;
    const int nos = A.nos;
    double s[3], fv[3], pr[3];
    load3( A.s, nos, i, s );
    load3( A.fv1, nos, i, fv );
    if( SOLVER == SOLVER_DEPONDT )
    {
        double axis[3]     = { 1 * fv[0], 1 * fv[1], 1 * fv[2] };
        const double angle = sqrt( dot3( fv, fv ) );
        normalize3( axis );
        rotate3( s, axis, angle, pr );
    }
    else
    {
        double t[3];
        neg_cross3( s, fv, t );
#pragma unroll
        for( int q = 0; q < 3; ++q )
        {
            pr[q] = 1 * s[q];
            pr[q] += 1 * t[q];
        }
        normalize3( pr );
    }
    store3( A.sp, nos, i, pr );
}

// adjoint of the final update and of FE2's virtual force
template<int SOLVER>
__global__ void __launch_bounds__( ADJ_THREADS ) k_adj_final( AdjArgs A )
{
    const int i = blockIdx.x * ADJ_THREADS + threadIdx.x;
    if( i >= A.nos )
        return;
    const int nos   = A.nos;
    const double mu = A.H.mu[basis_of( A.H, i )];
    double lam[3], s[3], sp[3], fv1[3], fv2[3], F2[3];
    load3( A.lam, nos, i, lam );
    load3( A.s, nos, i, s );
    load3( A.sp, nos, i, sp );
    load3( A.fv1, nos, i, fv1 );
    load3( A.fv2, nos, i, fv2 );
    load3( A.F2, nos, i, F2 );
    double ms[3], mfv1[3], mfv2[3], msp[3] = { 0, 0, 0 };
    if( SOLVER == SOLVER_DEPONDT )
    {
        double h[3];
#pragma unroll
        for( int q = 0; q < 3; ++q )
        {
            h[q] = 0.5 * fv1[q];
            h[q] += 0.5 * fv2[q];
        }
        const RotC r = rot_coeffs( sqrt( dot3( h, h ) ) );
        rot_T_v( h, r, lam, ms );
        double mh[3];
        rot_T_h( h, r, s, lam, mh );
#pragma unroll
        for( int q = 0; q < 3; ++q )
            mfv1[q] = mfv2[q] = 0.5 * mh[q];
    }
    else
    {
        double t1[3], t2[3], u[3];
        neg_cross3( s, fv1, t1 );
        neg_cross3( sp, fv2, t2 );
#pragma unroll
        for( int q = 0; q < 3; ++q )
        {
            u[q] = t1[q] * 0.5;
            u[q] += 1 * s[q];
            u[q] += 0.5 * t2[q];
        }
        double mu2[3];
        norm_d( u, lam, mu2 );
        double mt[3] = { 0.5 * mu2[0], 0.5 * mu2[1], 0.5 * mu2[2] };
        // t2 = -(s' x fv2): dL/ds' = mt x fv2, dL/dfv2 = s' x mt; t1 likewise with s, fv1
        cross3( mt, fv2, msp );
        cross3( sp, mt, mfv2 );
        double a1[3];
        cross3( mt, fv1, a1 );
        cross3( s, mt, mfv1 );
#pragma unroll
        for( int q = 0; q < 3; ++q )
            ms[q] = mu2[q] + a1[q];
    }
    // FE2: fv2 = k (F2 + alpha s' x F2) + xi + alpha s' x xi
    double xi[3], mF2[3];
    thermal_xi( A.L, mu, A.step, i, xi );
    vf_adjoint( vf_scale( A.L, mu ), A.L.damping, sp, F2, xi, mfv2, mF2, msp );
    const double mg2[3] = { -1 * mF2[0], -1 * mF2[1], -1 * mF2[2] };
    store3( A.mu_s, nos, i, ms );
    store3( A.mu_fv1, nos, i, mfv1 );
    store3( A.mu_sp, nos, i, msp );
    store3( A.mu_g, nos, i, mg2 );
}

// adjoint of the predictor and of FE1's virtual force (after mu_sp has its Hessian term)
template<int SOLVER>
__global__ void __launch_bounds__( ADJ_THREADS ) k_adj_predictor_back( AdjArgs A )
{
    const int i = blockIdx.x * ADJ_THREADS + threadIdx.x;
    if( i >= A.nos )
        return;
    const int nos   = A.nos;
    const double mu = A.H.mu[basis_of( A.H, i )];
    double s[3], sp[3], fv1[3], F1[3], msp[3], ms[3], mfv1[3];
    load3( A.s, nos, i, s );
    load3( A.sp, nos, i, sp );
    load3( A.fv1, nos, i, fv1 );
    load3( A.F1, nos, i, F1 );
    load3( A.mu_sp, nos, i, msp );
    load3( A.mu_s, nos, i, ms );
    load3( A.mu_fv1, nos, i, mfv1 );
    if( SOLVER == SOLVER_DEPONDT )
    {
        const RotC r = rot_coeffs( sqrt( dot3( fv1, fv1 ) ) );
        double a[3], b[3];
        rot_T_v( fv1, r, msp, a );
        rot_T_h( fv1, r, s, msp, b );
#pragma unroll
        for( int q = 0; q < 3; ++q )
        {
            ms[q] += a[q];
            mfv1[q] += b[q];
        }
    }
    else
    {
        double t1[3], u[3];
        neg_cross3( s, fv1, t1 );
#pragma unroll
        for( int q = 0; q < 3; ++q )
        {
            u[q] = 1 * s[q];
            u[q] += 1 * t1[q];
        }
        double mu1[3], a1[3], b1[3];
        norm_d( u, msp, mu1 );
        cross3( mu1, fv1, a1 );
        cross3( s, mu1, b1 );
#pragma unroll
        for( int q = 0; q < 3; ++q )
        {
            ms[q] += mu1[q] + a1[q];
            mfv1[q] += b1[q];
        }
    }
    double xi[3], mF1[3];
    thermal_xi( A.L, mu, A.step, i, xi );
    vf_adjoint( vf_scale( A.L, mu ), A.L.damping, s, F1, xi, mfv1, mF1, ms );
    const double mg1[3] = { -1 * mF1[0], -1 * mF1[1], -1 * mF1[2] };
    store3( A.mu_s, nos, i, ms );
    store3( A.mu_g, nos, i, mg1 );
}

// acc += H mu_g (Hessian of the energy at xg), and the parameter partials
// mu_g . d grad / d(J_shell, D_shell, B, K_basis, K_i) at xg.
__global__ void __launch_bounds__( ADJ_THREADS ) k_adj_hessian( AdjArgs A )
{
    const int i     = blockIdx.x * ADJ_THREADS + threadIdx.x;
    const bool act  = i < A.nos;
    const int nos   = A.nos;
    const HamDev & H = A.H;
    double pg[ADJ_NPG];
#pragma unroll
    for( int k = 0; k < ADJ_NPG; ++k )
        pg[k] = 0;
    if( act )
    {
        double hv[3], e;
        local_gradient( A.Hh, A.mu_g, nos, i, A.hv, hv, e );
        double v[3], x[3];
        load3( A.mu_g, nos, i, v );
        load3( A.xg, nos, i, x );
        const int ib   = basis_of( H, i );
        const int cell = i / H.nb;
        const int a    = cell % H.Na;
        const int rest = cell / H.Na;
        const int b    = rest % H.Nb;
        const int c    = rest / H.Nb;
        // cubic anisotropy: grad -2K s_q^3, Hessian -6K s_q^2 on the diagonal
        for( int k = 0; k < H.n_cubic; ++k )
            if( H.cubic_idx[k] == ib )
#pragma unroll
                for( int q = 0; q < 3; ++q )
                    hv[q] += -6.0 * H.cubic_K[k] * x[q] * x[q] * v[q];
        double acc[3];
        load3( A.acc, nos, i, acc );
#pragma unroll
        for( int q = 0; q < 3; ++q )
            acc[q] += hv[q];
        store3( A.acc, nos, i, acc );
        // exchange: grad_i -= J s_j
        for( int k = 0; k < H.n_exch; ++k )
        {
            const PairDev & p = H.exch[k];
            if( p.i != ib || p.pad >= ADJ_MAX_SHELLS )
                continue;
            const int j = pair_neighbour( H, a, b, c, p );
            if( j >= 0 )
            {
                const double xj[3] = { A.xg[j], A.xg[nos + j], A.xg[2 * nos + j] };
                const double val   = -dot3( v, xj );
                // static slot indices keep the accumulators in registers
#pragma unroll
                for( int sh = 0; sh < ADJ_MAX_SHELLS; ++sh )
                    if( p.pad == sh )
                        pg[ADJ_PG_J + sh] += val;
            }
        }
        // DMI: grad_i -= D (s_j x n)
        for( int k = 0; k < H.n_dmi; ++k )
        {
            const PairDev & p = H.dmi[k];
            if( p.i != ib || p.pad >= ADJ_MAX_SHELLS )
                continue;
            const int j = pair_neighbour( H, a, b, c, p );
            if( j >= 0 )
            {
                const double xj[3] = { A.xg[j], A.xg[nos + j], A.xg[2 * nos + j] };
                double cr[3];
                cross3( xj, p.n, cr );
                const double val = -dot3( v, cr );
#pragma unroll
                for( int sh = 0; sh < ADJ_MAX_SHELLS; ++sh )
                    if( p.pad == sh )
                        pg[ADJ_PG_D + sh] += val;
            }
        }
        // Zeeman: grad_i -= mu_i (B mu_B) n  ->  d/dB [Tesla] = -mu_i mu_B n
        pg[ADJ_PG_B] -= H.mu[ib] * mu_B * dot3( v, H.Bn );
        // uniaxial anisotropy per basis atom and per site: grad_i -= 2 K n (n.s_i)
#pragma unroll
        for( int k = 0; k < MAX_BASIS; ++k )
            if( k < H.n_aniso && H.aniso_idx[k] == ib )
                pg[ADJ_PG_K + k] -= 2.0 * dot3( v, H.aniso_n[k] ) * dot3( H.aniso_n[k], x );
        if( A.dK )
            A.dK[i] -= 2.0 * dot3( v, A.site_n ) * dot3( A.site_n, x );
    }
    // block sums of the parameter partials, added to this block's accumulators
    __shared__ double sh[ADJ_NPG][ADJ_THREADS / 64];
#pragma unroll
    for( int k = 0; k < ADJ_NPG; ++k )
    {
        double t = pg[k];
        for( int off = 32; off > 0; off >>= 1 )
            t += __shfl_down( t, off, 64 );
        if( threadIdx.x % 64 == 0 )
            sh[k][threadIdx.x / 64] = t;
    }
    __syncthreads();
    if( threadIdx.x < ADJ_NPG )
    {
        double t = 0;
        for( int w = 0; w < ADJ_THREADS / 64; ++w )
            t += sh[threadIdx.x][w];
        A.pg[(long long)blockIdx.x * ADJ_NPG + threadIdx.x] += t;
    }
}

// ---------------------------------------------------------------------------
// tangent-linear stages
// ---------------------------------------------------------------------------

// force, virtual force and their tangents at (x, dx)
__device__ __forceinline__ void tl_force( const TlArgs & T, int i, const double * x, const double * dx, double * F,
                                          double * fv, double * dF, double * dfv )
{
    const int nos   = T.nos;
    const HamDev & H = T.H;
    const double mu = H.mu[basis_of( H, i )];
    double g[3], e, dg[3], gd[3];
    local_gradient( H, x, nos, i, ( T.h && H.has_ddi ) ? T.h : nullptr, g, e );
    local_gradient( T.Hh, dx, nos, i, ( T.hd && H.has_ddi ) ? T.hd : nullptr, dg, e );
    local_gradient( T.Hd, x, nos, i, nullptr, gd, e );
    double xi_[3], dxi[3];
    load3( x, nos, i, xi_ );
    load3( dx, nos, i, dxi );
    const int ib = basis_of( H, i );
    for( int k = 0; k < H.n_cubic; ++k )
        if( H.cubic_idx[k] == ib )
#pragma unroll
            for( int q = 0; q < 3; ++q )
                dg[q] += -6.0 * H.cubic_K[k] * xi_[q] * xi_[q] * dxi[q];
#pragma unroll
    for( int q = 0; q < 3; ++q )
    {
        F[q]  = -1 * g[q];
        dF[q] = -( dg[q] + gd[q] );
    }
    double th[3];
    thermal_xi( T.L, mu, T.step, i, th );
    virtual_force( T.L, SOLVER_DEPONDT, 1.0 / mu, xi_, F, th, fv );
    const double k = vf_scale( T.L, mu ), a = T.L.damping;
    double c1[3], c2[3], c3[3];
    cross3( dxi, F, c1 );
    cross3( xi_, dF, c2 );
    cross3( dxi, th, c3 );
#pragma unroll
    for( int q = 0; q < 3; ++q )
        dfv[q] = k * ( dF[q] + a * ( c1[q] + c2[q] ) ) + a * c3[q];
}

template<int SOLVER, int STAGE>
__global__ void __launch_bounds__( ADJ_THREADS ) k_tl_stage( TlArgs T )
{
    const int i = blockIdx.x * ADJ_THREADS + threadIdx.x;
    if( i >= T.nos )
        return;
    const int nos = T.nos;
    double s[3], ds[3];
    load3( T.s, nos, i, s );
    load3( T.ds, nos, i, ds );
    double F[3], fv[3], dF[3], dfv[3];
    if( STAGE == 0 )
    {
        tl_force( T, i, T.s, T.ds, F, fv, dF, dfv );
        store3( T.fv1, nos, i, fv );
        store3( T.dfv1, nos, i, dfv );
        double p[3], dp[3];
        if( SOLVER == SOLVER_DEPONDT )
        {
            double axis[3]     = { 1 * fv[0], 1 * fv[1], 1 * fv[2] };
            const double angle = sqrt( dot3( fv, fv ) );
            normalize3( axis );
            rotate3( s, axis, angle, p );
            const RotC r = rot_coeffs( angle );
            double a1[3], a2[3];
            rot_v( fv, r, ds, a1 );
            rot_d_h( fv, r, s, dfv, a2 );
#pragma unroll
            for( int q = 0; q < 3; ++q )
                dp[q] = a1[q] + a2[q];
        }
        else
        {
            double t[3], u[3], dt1[3], dt2[3], du[3];
            neg_cross3( s, fv, t );
#pragma unroll
            for( int q = 0; q < 3; ++q )
            {
                u[q] = 1 * s[q];
                u[q] += 1 * t[q];
            }
            // dt = -(ds x fv) - (s x dfv)
            neg_cross3( ds, fv, dt1 );
            neg_cross3( s, dfv, dt2 );
#pragma unroll
            for( int q = 0; q < 3; ++q )
                du[q] = ds[q] + dt1[q] + dt2[q];
            norm_d( u, du, dp );
#pragma unroll
            for( int q = 0; q < 3; ++q )
                p[q] = u[q];
            normalize3( p );
        }
        store3( T.sp, nos, i, p );
        store3( T.dsp, nos, i, dp );
    }
    else
    {
        double sp[3], dsp[3], fv1[3], dfv1[3];
        load3( T.sp, nos, i, sp );
        load3( T.dsp, nos, i, dsp );
        load3( T.fv1, nos, i, fv1 );
        load3( T.dfv1, nos, i, dfv1 );
        tl_force( T, i, T.sp, T.dsp, F, fv, dF, dfv );
        double o[3], d_o[3];
        if( SOLVER == SOLVER_DEPONDT )
        {
            double h[3], dh[3];
#pragma unroll
            for( int q = 0; q < 3; ++q )
            {
                h[q] = 0.5 * fv1[q];
                h[q] += 0.5 * fv[q];
                dh[q] = 0.5 * dfv1[q] + 0.5 * dfv[q];
            }
            const double angle = sqrt( dot3( h, h ) );
            double axis[3]     = { h[0], h[1], h[2] };
            normalize3( axis );
            rotate3( s, axis, angle, o );
            const RotC r = rot_coeffs( angle );
            double a1[3], a2[3];
            rot_v( h, r, ds, a1 );
            rot_d_h( h, r, s, dh, a2 );
#pragma unroll
            for( int q = 0; q < 3; ++q )
                d_o[q] = a1[q] + a2[q];
        }
        else
        {
            double t1[3], t2[3], u[3], a1[3], a2[3], b1[3], b2[3], du[3];
            neg_cross3( s, fv1, t1 );
            neg_cross3( sp, fv, t2 );
#pragma unroll
            for( int q = 0; q < 3; ++q )
            {
                u[q] = t1[q] * 0.5;
                u[q] += 1 * s[q];
                u[q] += 0.5 * t2[q];
            }
            neg_cross3( ds, fv1, a1 );
            neg_cross3( s, dfv1, a2 );
            neg_cross3( dsp, fv, b1 );
            neg_cross3( sp, dfv, b2 );
#pragma unroll
            for( int q = 0; q < 3; ++q )
                du[q] = 0.5 * ( a1[q] + a2[q] ) + ds[q] + 0.5 * ( b1[q] + b2[q] );
            norm_d( u, du, d_o );
#pragma unroll
            for( int q = 0; q < 3; ++q )
                o[q] = u[q];
            normalize3( o );
        }
        store3( T.s_out, nos, i, o );
        store3( T.ds_out, nos, i, d_o );
    }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
int adj_n_blocks( int nos )
{
    return ( nos + ADJ_THREADS - 1 ) / ADJ_THREADS;
}

void adj_launch_predictor( const AdjArgs & A, hipStream_t st )
{
    const dim3 g( adj_n_blocks( A.nos ) ), b( ADJ_THREADS );
    if( A.solver == SOLVER_DEPONDT )
        hipLaunchKernelGGL( k_adj_predictor<SOLVER_DEPONDT>, g, b, 0, st, A );
    else
        hipLaunchKernelGGL( k_adj_predictor<SOLVER_HEUN>, g, b, 0, st, A );
}

void adj_launch_final( const AdjArgs & A, hipStream_t st )
{
    const dim3 g( adj_n_blocks( A.nos ) ), b( ADJ_THREADS );
    if( A.solver == SOLVER_DEPONDT )
        hipLaunchKernelGGL( k_adj_final<SOLVER_DEPONDT>, g, b, 0, st, A );
    else
        hipLaunchKernelGGL( k_adj_final<SOLVER_HEUN>, g, b, 0, st, A );
}

void adj_launch_hessian( const AdjArgs & A, hipStream_t st )
{
    hipLaunchKernelGGL( k_adj_hessian, dim3( adj_n_blocks( A.nos ) ), dim3( ADJ_THREADS ), 0, st, A );
}

void adj_launch_predictor_back( const AdjArgs & A, hipStream_t st )
{
    const dim3 g( adj_n_blocks( A.nos ) ), b( ADJ_THREADS );
    if( A.solver == SOLVER_DEPONDT )
        hipLaunchKernelGGL( k_adj_predictor_back<SOLVER_DEPONDT>, g, b, 0, st, A );
    else
        hipLaunchKernelGGL( k_adj_predictor_back<SOLVER_HEUN>, g, b, 0, st, A );
}

void tl_launch_stage( int stage, const TlArgs & T, hipStream_t st )
{
    const dim3 g( adj_n_blocks( T.nos ) ), b( ADJ_THREADS );
    if( T.solver == SOLVER_DEPONDT )
    {
        if( stage == 0 )
            hipLaunchKernelGGL( ( k_tl_stage<SOLVER_DEPONDT, 0> ), g, b, 0, st, T );
        else
            hipLaunchKernelGGL( ( k_tl_stage<SOLVER_DEPONDT, 1> ), g, b, 0, st, T );
    }
    else
    {
        if( stage == 0 )
            hipLaunchKernelGGL( ( k_tl_stage<SOLVER_HEUN, 0> ), g, b, 0, st, T );
        else
            hipLaunchKernelGGL( ( k_tl_stage<SOLVER_HEUN, 1> ), g, b, 0, st, T );
    }
}

} // namespace spirit_hip

namespace spirit_hip
{

// AoS [x0,y0,z0,x1,...] <-> SoA [x...][y...][z...] of one field, on the device
__global__ void k_aos_to_soa( const double * __restrict__ aos, double * __restrict__ soa, long long n )
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if( i >= n )
        return;
    soa[i]         = aos[3 * i];
    soa[n + i]     = aos[3 * i + 1];
    soa[2 * n + i] = aos[3 * i + 2];
}

__global__ void k_soa_to_aos( const double * __restrict__ soa, double * __restrict__ aos, long long n )
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if( i >= n )
        return;
    aos[3 * i]     = soa[i];
    aos[3 * i + 1] = soa[n + i];
    aos[3 * i + 2] = soa[2 * n + i];
}

void device_aos_to_soa( const double * aos, double * soa, long long n, hipStream_t st )
{
    hipLaunchKernelGGL( k_aos_to_soa, dim3( (unsigned)( ( n + 255 ) / 256 ) ), dim3( 256 ), 0, st, aos, soa, n );
}

void device_soa_to_aos( const double * soa, double * aos, long long n, hipStream_t st )
{
    hipLaunchKernelGGL( k_soa_to_aos, dim3( (unsigned)( ( n + 255 ) / 256 ) ), dim3( 256 ), 0, st, soa, aos, n );
}

} // namespace spirit_hip
