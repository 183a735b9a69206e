// C API: the GNEB method on the chain (Simulation_GNEB_Start, core/src/Spirit/Simulation.cpp:219-330)
// and the chain data its hook publishes (Method_GNEB::Hook_Post_Iteration, Method_GNEB.cpp:411-456).
#include "api_common.hpp"

#include <chrono>
#include <cmath>

using namespace spirit_hip;

namespace spirit_hip
{

// Cubic_Hermite_Spline::Interpolate, core/src/utility/Cubic_Hermite_Spline.cpp:10-47
static void hermite_interpolate( const std::vector<double> & x, const std::vector<double> & p,
                                 const std::vector<double> & m, int n_interp, std::vector<double> & ox,
                                 std::vector<double> & op )
{
    const size_t n    = p.size();
    const size_t npts = n + ( n - 1 ) * n_interp;
    ox.assign( npts, 0.0 );
    op.assign( npts, 0.0 );
    for( size_t i = 0; i + 1 < n; ++i )
    {
        const double x0 = x[i], x1 = x[i + 1], p0 = p[i], p1 = p[i + 1], m0 = m[i], m1 = m[i + 1];
        for( int j = 0; j < n_interp + 1; ++j )
        {
            const double t   = j / static_cast<double>( n_interp + 1 );
            const double h00 = 2 * std::pow( t, 3 ) - 3 * std::pow( t, 2 ) + 1;
            const double h10 = -2 * std::pow( t, 3 ) + 3 * std::pow( t, 2 );
            const double h01 = std::pow( t, 3 ) - 2 * std::pow( t, 2 ) + t;
            const double h11 = std::pow( t, 3 ) - std::pow( t, 2 );
            const size_t idx = i * ( n_interp + 1 ) + j;
            ox[idx]          = x0 + t * ( x1 - x0 );
            op[idx]          = h00 * p0 + h10 * p1 + h01 * m0 * ( x0 - x1 ) + h11 * m1 * ( x0 - x1 );
        }
    }
    ox.back() = x.back();
    op.back() = p.back();
}

// publish the chain data of the last iteration (energies, Rx, interpolation, torques)
static void gneb_update_chain( State * state, MethodGNEB & m )
{
    auto d   = m.runner->chain_data();
    auto & c = *state->chain;
    c.Rx     = d.Rx;
    for( int i = 0; i < c.noi() && i < (int)d.E.size(); ++i )
        c.images[i]->E = d.E[i];
    m.max_torque_all = d.torque;
    m.max_torque     = d.max_torque;
    hermite_interpolate( d.Rx, d.E, d.dE_dRx, c.gneb->n_E_interpolations, c.Rx_interpolated, c.E_interpolated );
}

void gneb_sync_host( State * state )
{
    auto m = state->method_chain;
    if( !m || !m->runner || !m->device_newer )
        return;
    m->runner->download_spins();
    auto & c = *state->chain;
    for( int i = 0; i < c.noi(); ++i )
        m->runner->download_effective_field( i, c.images[i]->eff_field.data() );
    m->device_newer = false;
}

void gneb_finalize( State * state )
{
    auto m = state->method_chain;
    if( m && m->runner )
    {
        m->device_newer = true;
        gneb_sync_host( state );
        gneb_update_chain( state, *m );
        m->runner.reset();
    }
    state->chain->iteration_allowed  = false;
    state->chain->singleshot_allowed = false;
}

static bool gneb_continue( State * state, const MethodGNEB & m )
{
    auto & c = *state->chain;
    return m.iteration < m.n_iterations && c.iteration_allowed && !stop_file_present()
           && !( m.max_torque < c.gneb->force_convergence );
}

static bool gneb_walltime_expired( State * state, const MethodGNEB & m )
{
    if( state->chain->gneb->max_walltime_sec <= 0 )
        return false;
    double s = std::chrono::duration<double>( std::chrono::steady_clock::now() - m.t_start ).count();
    return s > (double)state->chain->gneb->max_walltime_sec;
}

// n iterations on the device; the chain data and the stop reason are read back afterwards
static void gneb_advance( State * state, MethodGNEB & m, long n, bool check )
{
    auto t0    = std::chrono::steady_clock::now();
    long done  = m.runner->run( n, check );
    double sec = std::chrono::duration<double>( std::chrono::steady_clock::now() - t0 ).count();
    if( sec > 0 && done > 0 )
        m.ips = done / sec;
    m.iteration += done;
    m.device_newer = true;
    gneb_update_chain( state, m );
    if( m.runner->status() == 2 )
    {
        log_message( 2, "The geodesic distance between two images is zero! Stopping..." );
        state->chain->iteration_allowed = false;
    }
}

void gneb_n_shot( State * state, int N )
{
    auto m = state->method_chain;
    if( !m || !m->runner )
        throw api_error( 3, "Tried to use Simulation_SingleShot on the chain but no SingleShot simulation has been "
                            "started." );
    if( gneb_continue( state, *m ) && !gneb_walltime_expired( state, *m ) )
    {
        std::lock_guard<std::recursive_mutex> lk( state->chain->mtx );
        gneb_advance( state, *m, std::max( 0, N ), false );
        if( state->host_sync )
            gneb_sync_host( state );
    }
    if( !gneb_continue( state, *m ) || gneb_walltime_expired( state, *m ) )
        gneb_finalize( state );
}

// the chain's images as the runner sees them
static std::vector<ImageSpec> chain_specs( State * state )
{
    std::vector<ImageSpec> specs;
    for( auto & img : state->chain->images )
        specs.push_back( ImageSpec{ img->geometry, img->hamiltonian, *img->llg, img->spins.data() } );
    return specs;
}

// parameters changed on the host while a single-shot GNEB holds device copies
void gneb_notify_changed( State * state )
{
    try
    {
        auto m = state->method_chain;
        if( !m || !m->runner )
            return;
        m->runner->set_types( state->chain->image_type );
        m->runner->set_parameters( *state->chain->gneb );
    }
    catch( const std::exception & e )
    {
        log_message( 2, std::string( "GNEB parameter update: " ) + e.what() );
    }
}

} // namespace spirit_hip

extern "C" void Simulation_GNEB_Start( State * state, int solver_type, int n_iterations, int n_iterations_log,
                                       bool singleshot, Simulation_Run_Info * info, int idx_chain ) noexcept
try
{
    int idx_image = -1;
    std::shared_ptr<Image> image;
    std::shared_ptr<Chain> chain;
    from_indices( state, idx_image, idx_chain, image, chain );
    if( image->iteration_allowed || chain->iteration_allowed )
        throw api_error( 3, "Tried to use Simulation_Start on image -1 of chain " + std::to_string( idx_chain )
                                + ", but there is already a simulation running." );
    for( auto & img : chain->images )
        if( img->iteration_allowed )
        {
            log_message( 2, "There are still one or more simulations running on the specified chain! Please stop "
                            "them before starting a GNEB calculation." );
            return;
        }
    if( chain->noi() < 2 )
    {
        log_message( 2, "There are less than 2 images in the specified chain! Please insert more before starting a "
                        "GNEB calculation." );
        return;
    }
    sync_all_hosts( state );
    auto m = std::make_shared<MethodGNEB>();
    {
        std::lock_guard<std::recursive_mutex> lk( chain->mtx );
        chain->iteration_allowed  = true;
        chain->singleshot_allowed = singleshot;
        auto & gp                 = *chain->gneb;
        if( n_iterations > 0 )
            gp.n_iterations = n_iterations;
        if( n_iterations_log > 0 )
            gp.n_iterations_log = n_iterations_log;
        m->solver           = solver_type;
        m->n_iterations     = std::max( 1L, gp.n_iterations );
        m->n_iterations_log = gp.n_iterations_log;
        const char * names[] = { "VP", "SIB", "Depondt", "Heun", "RK4", "LBFGS_OSO", "LBFGS_Atlas", "VP_OSO" };
        m->solver_name       = ( solver_type >= 0 && solver_type <= 7 ) ? names[solver_type] : "unknown";
        try
        {
            chain->image_type.resize( chain->noi(), GNEB_NORMAL );
            m->runner.reset( new GNEBRunner( chain_specs( state ), gp, chain->image_type, solver_type ) );
        }
        catch( ... )
        {
            chain->iteration_allowed  = false;
            chain->singleshot_allowed = false;
            throw;
        }
        // not converged before the first iteration (Method_GNEB.cpp:64)
        m->max_torque = gp.force_convergence + 1.0;
        m->max_torque_all.assign( chain->noi(), 0.0 );
        m->t_start = std::chrono::steady_clock::now();
    }
    state->method_chain = m;
    if( singleshot )
        return;

    // Method::Iterate, Method.cpp:57-113 (n_iterations_amortize = 1): convergence is tested after
    // every iteration on the device; the host checks the stop requests between batches.
    while( gneb_continue( state, *m ) && !gneb_walltime_expired( state, *m ) )
    {
        long left = m->n_iterations - m->iteration;
        gneb_advance( state, *m, std::min( left, 500L ), true );
    }
    gneb_finalize( state );
    if( info )
    {
        info->max_torque       = (float)m->max_torque;
        info->total_iterations = (int)m->iteration;
        double ms = std::chrono::duration<double, std::milli>( std::chrono::steady_clock::now() - m->t_start ).count();
        info->total_walltime = (int)ms;
        info->total_ips      = ms > 0 ? float( m->iteration / ms * 1000.0 ) : 0.f;
    }
}
catch( ... )
{
    handle_exception( -1, idx_chain );
}

extern "C" void Chain_Update_Data( State * state, int idx_chain ) noexcept
try
{
    sync_all_hosts( state );
}
catch( ... )
{
    handle_exception( -1, idx_chain );
}
