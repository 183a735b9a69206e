// Differentiable LLG run: forward with checkpoints, discrete adjoint, tangent-linear run.
#include "adjoint.hpp"
#include "gpu_engine.hpp"

#include <algorithm>

namespace spirit_hip
{

AdjointRunner::AdjointRunner( const Geometry & g, const Hamiltonian & h, const LLGParameters & p, int solver_,
                              int n_steps_, int checkpoint_every )
        : geom( std::make_shared<Geometry>( g ) ),
          ham( std::make_shared<Hamiltonian>( h ) ),
          llg( p ),
          solver( solver_ ),
          n_steps( n_steps_ ),
          ckpt( checkpoint_every )
{
    if( solver != SOLVER_DEPONDT && solver != SOLVER_HEUN )
        throw std::runtime_error( "adjoint: Depondt or Heun only" );
    if( llg.direct_minimization )
        throw std::runtime_error( "adjoint: LLG dynamics only (direct minimisation is not differentiated)" );
    if( !h.quads.empty() )
        throw std::runtime_error( "adjoint: quadruplet interactions are not differentiated" );
    if( llg.stt_magnitude > 0 )
        throw std::runtime_error( "adjoint: spin-transfer torque is not differentiated" );
    if( llg.temperature_gradient_inclination != 0 )
        throw std::runtime_error( "adjoint: temperature gradients are not differentiated" );
    if( n_steps < 1 )
        throw std::runtime_error( "adjoint: at least one step" );
    if( ckpt < 1 || ckpt > n_steps )
        ckpt = n_steps;
    // one parameter per shell (or per input pair of a pair-list Hamiltonian)
    n_exch_shells = ham->exchange_shell_J.empty() ? (int)ham->exch_J_in.size() : (int)ham->exchange_shell_J.size();
    n_dmi_sh      = ham->dmi_shell_D.empty() ? (int)ham->dmi_D_in.size() : (int)ham->dmi_shell_D.size();
    if( n_exch_shells > ADJ_MAX_SHELLS || n_dmi_sh > ADJ_MAX_SHELLS )
        throw std::runtime_error( "adjoint: at most " + std::to_string( ADJ_MAX_SHELLS )
                                  + " exchange / DMI parameters" );
    const int nos = geom->nos;
    if( ham->site_K.empty() )
        ham->site_K.assign( nos, 0.0 ); // the per-site K_i are parameters even where they start at 0
    host_spins.assign( 3 * (size_t)nos, 0.0 );
    for( int i = 0; i < nos; ++i )
        host_spins[3 * i + 2] = 1.0;
    ImageSpec sp{ geom, ham, llg, host_spins.data() };
    run.reset( new LLGRunner( { sp }, solver ) );

    const size_t f3      = sizeof( double ) * 3 * (size_t)nos;
    const int n_ckpt     = ( n_steps + ckpt - 1 ) / ckpt;
    b_ckpt.resize( f3 * n_ckpt );
    b_tape_s.resize( f3 * ( ckpt + 1 ) );
    for( DeviceBuffer * b : { &b_F1, &b_fv1, &b_F2, &b_fv2 } )
        b->resize( f3 * ckpt );
    for( DeviceBuffer * b : { &b_sp, &b_lam, &b_ms, &b_msp, &b_mfv1, &b_mg, &b_hd, &b_dsp, &b_ds, &b_dsn } )
        b->resize( f3 );
    b_pg.resize( sizeof( double ) * ADJ_NPG * (size_t)adj_n_blocks( nos ) );
    upload_variants();
}

AdjointRunner::~AdjointRunner()
{
    if( run && run->stream )
        (void)hipStreamSynchronize( run->stream );
}

// the Hessian operator: the Hamiltonian without Zeeman and cubic terms (the energy is
// quadratic in the spins otherwise, so its gradient applied to v is H v)
void AdjointRunner::upload_variants()
{
    Hamiltonian hh    = *ham;
    hh.field_magnitude = 0;
    hh.cubic_idx.clear();
    hh.cubic_K.clear();
    hess_up.upload( *geom, hh );
}

void AdjointRunner::set_parameters( const double * J, const double * D, const double * B, const double * K,
                                    const double * K_site )
{
    auto & h = *ham;
    if( J )
    {
        if( !h.exchange_shell_J.empty() )
            for( int k = 0; k < n_exch_shells; ++k )
                h.exchange_shell_J[k] = J[k];
        else
            for( int k = 0; k < n_exch_shells; ++k )
                h.exch_J_in[k] = J[k];
        for( size_t p = 0; p < h.exch_J.size(); ++p )
            h.exch_J[p] = J[h.exch_shell[p]];
    }
    if( D )
    {
        if( !h.dmi_shell_D.empty() )
            for( int k = 0; k < n_dmi_sh; ++k )
                h.dmi_shell_D[k] = D[k];
        else
            for( int k = 0; k < n_dmi_sh; ++k )
                h.dmi_D_in[k] = D[k];
        for( size_t p = 0; p < h.dmi_D.size(); ++p )
            h.dmi_D[p] = D[h.dmi_shell[p]];
    }
    if( B )
        h.field_magnitude = B[0] * mu_B;
    if( K )
        for( size_t k = 0; k < h.aniso_K.size(); ++k )
            h.aniso_K[k] = K[k];
    if( K_site )
        std::copy( K_site, K_site + geom->nos, h.site_K.begin() );
    ++h.version; // the dipolar plan (ddi_version) is unaffected
    run->refresh_hamiltonian();
    upload_variants();
    tape_ok = false;
}

void AdjointRunner::get_parameters( double * J, double * D, double * B, double * K, double * K_site ) const
{
    const auto & h = *ham;
    if( J )
        for( int k = 0; k < n_exch_shells; ++k )
            J[k] = h.exchange_shell_J.empty() ? h.exch_J_in[k] : h.exchange_shell_J[k];
    if( D )
        for( int k = 0; k < n_dmi_sh; ++k )
            D[k] = h.dmi_shell_D.empty() ? h.dmi_D_in[k] : h.dmi_shell_D[k];
    if( B )
        B[0] = h.field_magnitude / mu_B;
    if( K )
        for( size_t k = 0; k < h.aniso_K.size(); ++k )
            K[k] = h.aniso_K[k];
    if( K_site )
        std::copy( h.site_K.begin(), h.site_K.end(), K_site );
}

void AdjointRunner::forward( const double * s0, double * sN )
{
    hipStream_t st  = run->stream;
    const size_t f3 = sizeof( double ) * 3 * (size_t)geom->nos;
    double * s      = run->dev_spins();
    HIPCHECK( hipMemcpyAsync( s, s0, f3, hipMemcpyDeviceToDevice, st ) );
    run->step_counter = 0;
    const int n_ckpt  = ( n_steps + ckpt - 1 ) / ckpt;
    for( int c = 0; c < n_ckpt; ++c )
    {
        HIPCHECK( hipMemcpyAsync( b_ckpt.as<char>() + f3 * c, s, f3, hipMemcpyDeviceToDevice, st ) );
        const int steps = std::min( ckpt, n_steps - c * ckpt );
        run->run( steps, 0, false, 0.0 );
    }
    HIPCHECK( hipMemcpyAsync( sN, s, f3, hipMemcpyDeviceToDevice, st ) );
    HIPCHECK( hipStreamSynchronize( st ) );
    tape_ok = true;
}

void AdjointRunner::backward( const double * dL_dsN, double * dL_ds0, double * dJ, double * dD, double * dB,
                              double * dK, double * dK_site )
{
    if( !tape_ok )
        throw std::runtime_error( "adjoint: run forward with the current parameters first" );
    hipStream_t st    = run->stream;
    const int nos     = geom->nos;
    const size_t n3   = 3 * (size_t)nos;
    const size_t f3   = sizeof( double ) * n3;
    const int n_blk   = adj_n_blocks( nos );
    HIPCHECK( hipMemsetAsync( b_pg.as<void>(), 0, b_pg.size(), st ) );
    DeviceBuffer dk( sizeof( double ) * nos );
    HIPCHECK( hipMemsetAsync( dk.as<void>(), 0, dk.size(), st ) );
    HIPCHECK( hipMemcpyAsync( b_lam.as<void>(), dL_dsN, f3, hipMemcpyDeviceToDevice, st ) );

    AdjArgs A{};
    A.H      = run->image_host( 0 ).ham;
    A.Hh     = hess_up.dev;
    A.L      = run->image_host( 0 ).llg;
    A.nos    = nos;
    A.solver = solver;
    for( int c = 0; c < 3; ++c )
        A.site_n[c] = ham->site_n[c];
    A.sp     = b_sp.as<double>();
    A.mu_sp  = b_msp.as<double>();
    A.mu_fv1 = b_mfv1.as<double>();
    A.mu_g   = b_mg.as<double>();
    A.dK     = dk.as<double>();
    A.pg     = b_pg.as<double>();
    double * lam = b_lam.as<double>();
    double * ms  = b_ms.as<double>();

    const int n_ckpt = ( n_steps + ckpt - 1 ) / ckpt;
    for( int seg = n_ckpt - 1; seg >= 0; --seg )
    {
        const int start = seg * ckpt;
        const int len   = std::min( ckpt, n_steps - start );
        double * ts     = b_tape_s.as<double>();
        HIPCHECK( hipMemcpyAsync( ts, b_ckpt.as<char>() + f3 * seg, f3, hipMemcpyDeviceToDevice, st ) );
        for( int j = 0; j < len; ++j )
            run->tape_step( (uint64_t)( start + j ), ts + n3 * j, ts + n3 * ( j + 1 ), b_F1.as<double>() + n3 * j,
                            b_fv1.as<double>() + n3 * j, b_F2.as<double>() + n3 * j, b_fv2.as<double>() + n3 * j );
        for( int j = len - 1; j >= 0; --j )
        {
            A.step = (uint64_t)( start + j );
            A.s    = ts + n3 * j;
            A.F1   = b_F1.as<double>() + n3 * j;
            A.fv1  = b_fv1.as<double>() + n3 * j;
            A.F2   = b_F2.as<double>() + n3 * j;
            A.fv2  = b_fv2.as<double>() + n3 * j;
            A.lam  = lam;
            A.mu_s = ms;
            adj_launch_predictor( A, st );
            adj_launch_final( A, st );
            run->compute_ddi_field( A.mu_g );
            A.hv  = run->dev_ddi_field();
            A.xg  = A.sp;
            A.acc = A.mu_sp;
            adj_launch_hessian( A, st );
            adj_launch_predictor_back( A, st );
            run->compute_ddi_field( A.mu_g );
            A.hv  = run->dev_ddi_field();
            A.xg  = A.s;
            A.acc = ms;
            adj_launch_hessian( A, st );
            std::swap( lam, ms );
        }
    }
    HIPCHECK( hipGetLastError() );
    HIPCHECK( hipMemcpyAsync( dL_ds0, lam, f3, hipMemcpyDeviceToDevice, st ) );
    std::vector<double> pg( (size_t)ADJ_NPG * n_blk );
    HIPCHECK( hipMemcpyAsync( pg.data(), b_pg.as<void>(), sizeof( double ) * pg.size(), hipMemcpyDeviceToHost, st ) );
    if( dK_site )
        HIPCHECK( hipMemcpyAsync( dK_site, dk.as<void>(), dk.size(), hipMemcpyDeviceToHost, st ) );
    HIPCHECK( hipStreamSynchronize( st ) );
    std::vector<double> tot( ADJ_NPG, 0.0 );
    for( int b = 0; b < n_blk; ++b )
        for( int k = 0; k < ADJ_NPG; ++k )
            tot[k] += pg[(size_t)b * ADJ_NPG + k];
    if( dJ )
        for( int k = 0; k < n_exch_shells; ++k )
            dJ[k] = tot[ADJ_PG_J + k];
    if( dD )
        for( int k = 0; k < n_dmi_sh; ++k )
            dD[k] = tot[ADJ_PG_D + k];
    if( dB )
        dB[0] = tot[ADJ_PG_B];
    if( dK )
        for( int k = 0; k < n_aniso(); ++k )
            dK[k] = tot[ADJ_PG_K + k];
}

void AdjointRunner::tangent( const double * s0, const double * ds0, const double * dJ, const double * dD,
                             const double * dB, const double * dK, const double * dK_site, double * sN, double * dsN )
{
    hipStream_t st  = run->stream;
    const int nos   = geom->nos;
    const size_t n3 = 3 * (size_t)nos;
    const size_t f3 = sizeof( double ) * n3;
    // parameter-direction Hamiltonian: the same terms with the parameter tangents, no dipoles
    Hamiltonian hd = *ham;
    for( size_t p = 0; p < hd.exch_J.size(); ++p )
        hd.exch_J[p] = dJ ? dJ[hd.exch_shell[p]] : 0.0;
    for( size_t p = 0; p < hd.dmi_D.size(); ++p )
        hd.dmi_D[p] = dD ? dD[hd.dmi_shell[p]] : 0.0;
    hd.field_magnitude = dB ? dB[0] * mu_B : 0.0;
    for( size_t k = 0; k < hd.aniso_K.size(); ++k )
        hd.aniso_K[k] = dK ? dK[k] : 0.0;
    if( dK_site )
        std::copy( dK_site, dK_site + nos, hd.site_K.begin() );
    else
        std::fill( hd.site_K.begin(), hd.site_K.end(), 0.0 );
    hd.cubic_idx.clear();
    hd.cubic_K.clear();
    hd.ddi_method = DDI_NONE;
    delta_up.upload( *geom, hd );

    TlArgs T{};
    T.H      = run->image_host( 0 ).ham;
    T.Hh     = hess_up.dev;
    T.Hd     = delta_up.dev;
    T.L      = run->image_host( 0 ).llg;
    T.nos    = nos;
    T.solver = solver;
    double * x   = b_tape_s.as<double>();
    double * x2  = b_tape_s.as<double>() + n3;
    double * dx  = b_ds.as<double>();
    double * dx2 = b_dsn.as<double>();
    HIPCHECK( hipMemcpyAsync( x, s0, f3, hipMemcpyDeviceToDevice, st ) );
    HIPCHECK( hipMemcpyAsync( dx, ds0, f3, hipMemcpyDeviceToDevice, st ) );
    T.sp   = b_sp.as<double>();
    T.dsp  = b_dsp.as<double>();
    T.fv1  = b_fv1.as<double>();
    T.dfv1 = b_mfv1.as<double>();
    T.hd   = run->dev_ddi_field() ? b_hd.as<double>() : nullptr;
    for( int n = 0; n < n_steps; ++n )
    {
        T.step = (uint64_t)n;
        T.s    = x;
        T.ds   = dx;
        for( int stage = 0; stage < 2; ++stage )
        {
            const double * cx  = stage == 0 ? x : T.sp;
            const double * cdx = stage == 0 ? dx : T.dsp;
            if( run->dev_ddi_field() )
            {
                run->compute_ddi_field( cdx );
                HIPCHECK( hipMemcpyAsync( b_hd.as<void>(), run->dev_ddi_field(), f3, hipMemcpyDeviceToDevice, st ) );
                run->compute_ddi_field( cx );
            }
            T.h      = run->dev_ddi_field();
            T.s_out  = x2;
            T.ds_out = dx2;
            tl_launch_stage( stage, T, st );
        }
        std::swap( x, x2 );
        std::swap( dx, dx2 );
    }
    HIPCHECK( hipGetLastError() );
    HIPCHECK( hipMemcpyAsync( sN, x, f3, hipMemcpyDeviceToDevice, st ) );
    HIPCHECK( hipMemcpyAsync( dsN, dx, f3, hipMemcpyDeviceToDevice, st ) );
    HIPCHECK( hipStreamSynchronize( st ) );
}

} // namespace spirit_hip
