// GNEB chain runner: the iteration sequence of Method_GNEB with the VP, SIB, Depondt, Heun,
// RK4, VP_OSO, LBFGS_OSO and LBFGS_Atlas solvers, enqueued on the device (core/src/engine/Method_GNEB.cpp,
// core/include/engine/Solver_*.hpp, core/src/engine/Method.cpp:57-113).
#include "gpu_engine.hpp"

#include <cstring>

namespace spirit_hip
{

GNEBRunner::GNEBRunner( const std::vector<ImageSpec> & images, const GNEBParameters & par,
                        const std::vector<int> & types, int solver_ )
        : specs( images ), solver( solver_ )
{
    if( solver != SOLVER_VP && solver != SOLVER_SIB && solver != SOLVER_DEPONDT && solver != SOLVER_HEUN
        && solver != SOLVER_RK4 && solver != SOLVER_VP_OSO && solver != SOLVER_LBFGS_OSO
        && solver != SOLVER_LBFGS_ATLAS )
        throw std::runtime_error( "GNEB solver " + std::to_string( solver ) + " is not available in the HIP engine" );
    noi = (int)images.size();
    if( noi < 2 )
        throw std::runtime_error( "GNEB needs at least 2 images" );
    if( noi > 256 )
        throw std::runtime_error( "GNEB in the HIP engine supports at most 256 images" );
    nos = images[0].geometry->nos;
    for( auto & sp : images )
        if( sp.geometry->nos != nos )
            throw std::runtime_error( "GNEB needs images with equal numbers of spins" );
    field.reset( new LLGRunner( images, SOLVER_DEPONDT ) );
    bpi = ( nos + GNEB_THREADS - 1 ) / GNEB_THREADS;

    const size_t f3 = sizeof( double ) * 3 * (size_t)nos * noi;
    hipStream_t st  = field->stream;
    for( DeviceBuffer * b : { &b_eff, &b_Fg, &b_tau, &b_Ft, &b_fv, &b_v, &b_fprev } )
    {
        b->resize( f3 );
        HIPCHECK( hipMemsetAsync( b->as<void>(), 0, f3, st ) );
    }
    if( solver != SOLVER_VP )
        for( DeviceBuffer * b : { &b_fv0, &b_pred, &b_tmp } )
        {
            b->resize( f3 );
            HIPCHECK( hipMemsetAsync( b->as<void>(), 0, f3, st ) );
        }
    b_ftr.resize( sizeof( double ) * 6 * (size_t)nos );
    HIPCHECK( hipMemsetAsync( b_ftr.as<void>(), 0, b_ftr.size(), st ) );
    b_part.resize( sizeof( double ) * GNEB_NPART * (size_t)noi * bpi );
    HIPCHECK( hipMemsetAsync( b_part.as<void>(), 0, b_part.size(), st ) );
    b_ch.resize( sizeof( double ) * GNEB_NCH * (size_t)noi );
    HIPCHECK( hipMemsetAsync( b_ch.as<void>(), 0, b_ch.size(), st ) );
    b_types.resize( sizeof( int ) * noi );
    b_stop.resize( sizeof( int ) );
    b_iter.resize( sizeof( long long ) );
    b_count.resize( sizeof( int ) * GNEB_COUNT_STRIDE * ( 1 + (size_t)noi ) );
    HIPCHECK( hipMemsetAsync( b_count.as<void>(), 0, b_count.size(), st ) );
    b_isum.resize( sizeof( double ) * GNEB_ISUM * (size_t)noi );
    HIPCHECK( hipMemsetAsync( b_isum.as<void>(), 0, b_isum.size(), st ) );
    HIPCHECK( hipMemsetAsync( b_stop.as<void>(), 0, sizeof( int ), st ) );
    HIPCHECK( hipMemsetAsync( b_iter.as<void>(), 0, sizeof( long long ), st ) );

    A.imgs     = field->dev_images();
    A.noi      = noi;
    A.nos      = nos;
    A.bpi      = bpi;
    A.n_blocks = noi * bpi;
    A.xc       = field->dev_spins();
    A.eff      = b_eff.as<double>();
    A.Fg       = b_Fg.as<double>();
    A.tau      = b_tau.as<double>();
    A.Ft       = b_Ft.as<double>();
    A.ftr      = b_ftr.as<double>();
    A.fv       = b_fv.as<double>();
    A.v        = b_v.as<double>();
    A.fprev    = b_fprev.as<double>();
    A.part     = b_part.as<double>();
    A.ch       = b_ch.as<double>();
    A.types    = b_types.as<int>();
    A.stop     = b_stop.as<int>();
    A.iter     = b_iter.as<long long>();
    A.count    = b_count.as<int>();
    A.isum     = b_isum.as<double>();
    set_types( types );
    set_parameters( par );
    if( min_is_solver( solver ) && solver != SOLVER_VP )
        setup_minimiser();
    HIPCHECK( hipStreamSynchronize( st ) );
}

// State of VP_OSO / LBFGS_OSO / LBFGS_Atlas over the chain, zero as after the solver's
// Initialize; the atlas charts from the signs of the images' initial s_z
// (Solver_LBFGS_Atlas.hpp:34-41).
void GNEBRunner::setup_minimiser()
{
    hipStream_t st   = field->stream;
    const size_t ns  = (size_t)noi * nos;
    const int dim    = solver == SOLVER_LBFGS_ATLAS ? 2 : 3;
    auto zeroed      = [&]( DeviceBuffer & b, size_t bytes ) {
        b.resize( bytes );
        HIPCHECK( hipMemsetAsync( b.as<void>(), 0, bytes, st ) );
    };
    const size_t f3 = sizeof( double ) * 3 * ns, fd = sizeof( double ) * dim * ns;
    zeroed( b_mv, f3 );
    zeroed( b_mprev, f3 );
    zeroed( b_mgrad, fd );
    zeroed( b_mgpr, fd );
    zeroed( b_msd, fd );
    zeroed( b_mq, fd );
    for( int k = 0; k < MIN_MEMORY; ++k )
    {
        zeroed( b_mda[k], fd );
        zeroed( b_mdg[k], fd );
    }
    zeroed( b_mpart, sizeof( double ) * MIN_SLOTS * (size_t)noi * bpi );
    zeroed( b_msc, sizeof( MinScalars ) );
    std::vector<double> a3( ns, 1.0 );
    for( int k = 0; k < noi; ++k )
        for( int i = 0; i < nos; ++i )
            a3[(size_t)k * nos + i] = specs[k].spins_aos[3 * (size_t)i + 2] > 0 ? 1.0 : -1.0;
    b_ma3.resize( sizeof( double ) * ns );
    HIPCHECK( hipMemcpyAsync( b_ma3.as<void>(), a3.data(), sizeof( double ) * ns, hipMemcpyHostToDevice, st ) );
    HIPCHECK( hipStreamSynchronize( st ) );
    MinArgs & M = min_args;
    M.v         = b_mv.as<double>();
    M.prev      = b_mprev.as<double>();
    M.f         = nullptr;
    M.grad      = b_mgrad.as<double>();
    M.grad_pr   = b_mgpr.as<double>();
    M.sd        = b_msd.as<double>();
    M.q         = b_mq.as<double>();
    for( int k = 0; k < MIN_MEMORY; ++k )
    {
        M.da[k] = b_mda[k].as<double>();
        M.dg[k] = b_mdg[k].as<double>();
    }
    M.a3       = b_ma3.as<double>();
    M.part     = b_mpart.as<double>();
    M.F_ext    = b_Ft.as<double>();
    M.sc       = b_msc.as<MinScalars>();
    M.n_blocks = noi * bpi;
    M.noi      = noi;
    M.bpi      = bpi;
    M.dim      = dim;
    M.maxmove  = solver == SOLVER_LBFGS_ATLAS ? 0.05 : Pi / 200.0; // Solver_LBFGS_*.hpp Initialize
}

GNEBRunner::~GNEBRunner()
{
    if( field && field->stream )
        (void)hipStreamSynchronize( field->stream );
}

void GNEBRunner::set_types( const std::vector<int> & types )
{
    std::vector<int> t( noi, 0 );
    for( int i = 0; i < noi && i < (int)types.size(); ++i )
        t[i] = types[i];
    HIPCHECK( hipMemcpy( b_types.as<void>(), t.data(), sizeof( int ) * noi, hipMemcpyHostToDevice ) );
}

void GNEBRunner::set_parameters( const GNEBParameters & par )
{
    GnebParamsDev & P            = A.P;
    P.m                          = 1.0;
    P.spring_constant            = par.spring_constant;
    P.spring_force_ratio         = par.spring_force_ratio;
    P.path_shortening_constant   = par.path_shortening_constant;
    P.delta_Rx_left              = par.equilibrium_delta_Rx_left;
    P.delta_Rx_right             = par.equilibrium_delta_Rx_right;
    P.force_convergence          = par.force_convergence;
    P.moving_endpoints           = par.moving_endpoints ? 1 : 0;
    P.translating_endpoints      = par.translating_endpoints ? 1 : 0;
    P.escape_first               = par.escape_first ? 1 : 0;
}

// Method_GNEB::Calculate_Force (+ _Virtual) of the configuration x, Method_GNEB.cpp:88-391
void GNEBRunner::calc_force( double * x, bool vp )
{
    hipStream_t st = field->stream;
    A.x            = x;
    field->compute_ddi_field( x );
    A.h = field->dev_ddi_field();
    gneb_launch_pass( GPASS_GRAD, A, st );
    gneb_launch_pass( GPASS_TANGENT, A, st );
    if( A.P.moving_endpoints && A.P.translating_endpoints )
    {
        gneb_launch_pass( GPASS_TRANSLATE, A, st );
    }
    if( A.P.path_shortening_constant > 0 )
    {
        gneb_launch_pass( GPASS_SHRINK1, A, st );
        gneb_launch_pass( GPASS_SHRINK2, A, st );
        gneb_launch_pass( GPASS_SHRINK3, A, st );
        gneb_launch_pass( GPASS_SHRINK4, A, st );
    }
    gneb_launch_pass( vp ? GPASS_FORCE_VP : GPASS_FORCE_FV, A, st );
}

// One solver iteration and Hook_Post_Iteration (Method_GNEB.cpp:411-456)
void GNEBRunner::iteration()
{
    hipStream_t st = field->stream;
    double * xs    = field->dev_spins();
    if( solver == SOLVER_VP )
    {
        calc_force( xs, true );
        A.x = xs;
        gneb_launch_pass( GPASS_VP_UPDATE, A, st );
        return;
    }
    if( min_is_solver( solver ) )
    {
        // VP_OSO, LBFGS_OSO, LBFGS_Atlas on F_total (Solver_VP_OSO.hpp:33-110,
        // Solver_LBFGS_OSO.hpp:37-72, Solver_LBFGS_Atlas.hpp:44-101)
        A.fv = b_fv.as<double>();
        calc_force( xs, false );
        StepArgs SA{};
        SA.imgs     = A.imgs;
        SA.s        = xs;
        SA.stop     = A.stop;
        SA.img_stop = A.stop;
        min_launch_chain_update( solver, SA, min_args, nos, st );
        A.x = xs;
        gneb_launch_pass( GPASS_HOOK, A, st );
        return;
    }
    GnebSolverArgs S{};
    S.noi  = noi;
    S.nos  = nos;
    S.bpi  = bpi;
    S.x    = xs;
    S.pred = b_pred.as<double>();
    S.tmp  = b_tmp.as<double>();
    S.fv0  = b_fv0.as<double>();
    S.stop = A.stop;
    // first force evaluation writes the first virtual force
    A.fv = b_fv0.as<double>();
    calc_force( xs, false );
    S.fv = b_fv0.as<double>();
    gneb_launch_solver_stage( solver, 0, S, st );
    A.fv = b_fv.as<double>();
    S.fv = b_fv.as<double>();
    const int n_more = solver == SOLVER_RK4 ? 3 : 1;
    for( int k = 1; k <= n_more; ++k )
    {
        calc_force( b_pred.as<double>(), false );
        gneb_launch_solver_stage( solver, k, S, st );
    }
    A.x = xs;
    gneb_launch_pass( GPASS_HOOK, A, st );
}

long GNEBRunner::run( long n, bool check_convergence )
{
    hipStream_t st          = field->stream;
    A.P.check_convergence = check_convergence ? 1 : 0;
    HIPCHECK( hipMemsetAsync( b_stop.as<void>(), 0, sizeof( int ), st ) );
    HIPCHECK( hipMemsetAsync( b_iter.as<void>(), 0, sizeof( long long ), st ) );
    for( long k = 0; k < n; ++k )
        iteration();
    HIPCHECK( hipGetLastError() );
    long long done = 0;
    HIPCHECK( hipMemcpyAsync( &done, b_iter.as<void>(), sizeof( long long ), hipMemcpyDeviceToHost, st ) );
    HIPCHECK( hipStreamSynchronize( st ) );
    return (long)done;
}

int GNEBRunner::status()
{
    int s = 0;
    HIPCHECK( hipStreamSynchronize( field->stream ) );
    HIPCHECK( hipMemcpy( &s, b_stop.as<void>(), sizeof( int ), hipMemcpyDeviceToHost ) );
    return s;
}

GNEBRunner::ChainData GNEBRunner::chain_data()
{
    std::vector<double> ch( (size_t)GNEB_NCH * noi );
    HIPCHECK( hipStreamSynchronize( field->stream ) );
    HIPCHECK( hipMemcpy( ch.data(), b_ch.as<void>(), sizeof( double ) * ch.size(), hipMemcpyDeviceToHost ) );
    ChainData d;
    for( int i = 0; i < noi; ++i )
    {
        const double * c = ch.data() + (size_t)i * GNEB_NCH;
        d.E.push_back( c[GC_E] );
        d.Rx.push_back( c[GC_RX] );
        d.dE_dRx.push_back( c[GC_DEDRX] );
        d.torque.push_back( c[GC_TORQUE] );
    }
    d.max_torque = ch[GC_MAXT];
    return d;
}

void GNEBRunner::download_spins()
{
    field->download_spins();
}

void GNEBRunner::upload_spins()
{
    field->upload_spins();
}

void GNEBRunner::download_effective_field( int img, double * aos )
{
    std::vector<double> soa( 3 * (size_t)nos );
    HIPCHECK( hipStreamSynchronize( field->stream ) );
    HIPCHECK( hipMemcpy( soa.data(), b_eff.as<double>() + 3 * (size_t)img * nos, sizeof( double ) * soa.size(),
                         hipMemcpyDeviceToHost ) );
    soa_to_aos( soa.data(), aos, nos );
}

} // namespace spirit_hip
