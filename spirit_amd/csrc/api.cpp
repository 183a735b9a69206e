// The C API of the engine (libSpirit.so): the reference's exported functions
// (core/include/Spirit/*.h) backed by the HIP runner.  Every entry point catches
// all exceptions, logs them and returns a default value, as the reference does
// (core/include/utility/Exception.hpp:97-127).
#include "api_common.hpp"
#include "mc.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <ctime>
#include <fstream>
#include <thread>

using namespace spirit_hip;

namespace spirit_hip
{

Image::Image( const Image & o )
        : geometry( o.geometry ),
          hamiltonian( std::make_shared<Hamiltonian>( *o.hamiltonian ) ),
          llg( std::make_shared<LLGParameters>( *o.llg ) ),
          mc( std::make_shared<MCParameters>( *o.mc ) ),
          prng( o.prng ),
          spins( o.spins ),
          eff_field( o.eff_field ),
          E( o.E ),
          E_array( o.E_array ),
          max_torque( o.max_torque ),
          Rx( o.Rx )
{
}

void from_indices( State * state, int & idx_image, int & idx_chain, std::shared_ptr<Image> & image,
                   std::shared_ptr<Chain> & chain )
{
    if( !state || !state->chain )
        throw api_error( 2, "The State pointer is invalid" );
    idx_chain = 0;
    chain     = state->chain;
    if( idx_image >= chain->noi() )
        throw api_error( 3, "Index " + std::to_string( idx_image ) + " points to non-existent image (NOI="
                                + std::to_string( chain->noi() ) + "). No action taken." );
    if( idx_image < 0 )
    {
        image     = state->active_image;
        idx_image = state->idx_active_image;
    }
    else
        image = chain->images[idx_image];
}

void handle_exception( int idx_image, int idx_chain )
{
    try
    {
        throw;
    }
    catch( const api_error & e )
    {
        log_message( e.level, e.what() );
    }
    catch( const std::exception & e )
    {
        log_message( 2, std::string( "API call failed (image " ) + std::to_string( idx_image ) + ", chain "
                            + std::to_string( idx_chain ) + "): " + e.what() );
    }
    catch( ... )
    {
        log_message( 2, "API call failed with an unknown exception" );
    }
}

std::shared_ptr<Image> make_image( const std::string & cfg )
{
    auto img         = std::make_shared<Image>();
    img->geometry    = geometry_from_config( cfg );
    img->hamiltonian = hamiltonian_from_config( cfg, *img->geometry );
    img->llg         = std::make_shared<LLGParameters>( llg_parameters_from_config( cfg ) );
    img->mc          = std::make_shared<MCParameters>( mc_parameters_from_config( cfg ) );
    img->prng        = std::mt19937( (unsigned)img->llg->seed );
    img->spins.assign( 3 * (size_t)img->geometry->nos, 0.0 );
    img->eff_field.assign( 3 * (size_t)img->geometry->nos, 0.0 );
    for( int i = 0; i < img->geometry->nos; ++i )
        img->spins[3 * i + 2] = 1.0;
    return img;
}

// Bring the host mirror of an image up to date with a single-shot method's device spins.
void sync_image_host( State * state, int idx_image )
{
    auto it = state->method_image.find( idx_image );
    if( it == state->method_image.end() || !it->second )
        return;
    auto & m = *it->second;
    if( m.runner && m.device_newer )
    {
        m.runner->download_spins();
        m.runner->download_effective_field( 0, m.image->eff_field.data() );
        m.device_newer = false;
    }
}

void sync_all_hosts( State * state )
{
    for( auto & kv : state->method_image )
        sync_image_host( state, kv.first );
    gneb_sync_host( state );
}

// Parameters or spins changed on the host while a single-shot method holds device copies.
void notify_image_changed( State * state, int idx_image, bool spins_changed, bool ham_changed, bool llg_changed )
{
    auto mc = state->method_mc.find( idx_image );
    if( mc != state->method_mc.end() )
    {
        // a single-shot Monte Carlo run keeps the host spins current after every N_Shot
        if( ham_changed )
            mc->second->runner->refresh_hamiltonian();
        if( spins_changed )
            mc->second->runner->upload_spins();
        return;
    }
    auto it = state->method_image.find( idx_image );
    if( it == state->method_image.end() || !it->second || !it->second->runner )
        return;
    auto & m = *it->second;
    if( ham_changed )
        m.runner->refresh_hamiltonian();
    if( llg_changed )
        m.runner->set_llg( 0, *m.image->llg );
    if( spins_changed )
    {
        m.runner->upload_spins();
        m.device_newer = false;
    }
}

bool stop_file_present()
{
    std::ifstream f( "STOP" );
    return f.good();
}

} // namespace spirit_hip

// ===========================================================================
// State.h
// ===========================================================================
extern "C" State * State_Setup( const char * config_file, bool quiet ) noexcept
try
{
    std::string cfg = config_file ? config_file : "";
    g_quiet         = quiet;
    if( !cfg.empty() )
    {
        ConfigFile f( cfg );
        if( f.ok() )
        {
            int lvl = g_log_console_level;
            if( f.read( lvl, "log_console_level" ) )
                g_log_console_level = lvl;
        }
        else
            log_message( 2, "Could not open config file \"" + cfg + "\"; using defaults" );
    }
    auto * state       = new State();
    state->config_file = cfg;
    state->quiet       = quiet;
    {
        std::time_t t = std::time( nullptr );
        char buf[64];
        std::strftime( buf, sizeof( buf ), "%Y-%m-%d_%H-%M-%S", std::localtime( &t ) );
        state->datetime = buf;
    }
    state->chain       = std::make_shared<Chain>();
    state->chain->gneb = std::make_shared<GNEBParameters>( gneb_parameters_from_config( cfg ) );
    auto img           = make_image( cfg );
    // Configurations::Random would need the image's prng; the reference sets a random state here
    // (State.cpp:125-130) -- the setup keeps +z and leaves the choice to the caller.
    state->chain->images.push_back( img );
    state->chain->image_type.push_back( GNEB_NORMAL );
    state->chain->Rx.assign( 1, 0.0 );
    state->active_image     = img;
    state->idx_active_image = 0;
    state->nos              = img->nos();
    state->noi              = 1;
    state->clipboard_image  = nullptr;
    // State_Setup seeds a random configuration, State.cpp:125-130
    {
        std::uniform_real_distribution<double> dist( -1, 1 );
        for( int i = 0; i < img->nos(); ++i )
        {
            double vz  = dist( img->prng );
            double phi = dist( img->prng ) * Pi;
            double rxy = std::sqrt( 1 - vz * vz );
            img->spins[3 * i]     = rxy * std::cos( phi );
            img->spins[3 * i + 1] = rxy * std::sin( phi );
            img->spins[3 * i + 2] = vz;
        }
    }
    return state;
}
catch( ... )
{
    handle_exception( -1, -1 );
    return nullptr;
}

extern "C" void State_Delete( State * state ) noexcept
try
{
    if( !state )
        return;
    state->method_image.clear();
    state->method_mc.clear();
    state->method_chain.reset();
    delete state;
}
catch( ... )
{
    handle_exception( -1, -1 );
}

extern "C" void State_Update( State * state ) noexcept
try
{
    if( !state )
        return;
    auto & c = *state->chain;
    state->noi = c.noi();
    if( state->idx_active_image >= c.noi() )
        state->idx_active_image = c.noi() - 1;
    state->active_image = c.images[state->idx_active_image];
    state->nos          = state->active_image->nos();
}
catch( ... )
{
    handle_exception( -1, -1 );
}

extern "C" void State_To_Config( State * state, const char * config_file, const char * comment ) noexcept
try
{
    auto & img = *state->active_image;
    auto & g   = *img.geometry;
    auto & h   = *img.hamiltonian;
    std::ofstream f( config_file );
    if( comment && *comment )
        f << "# " << comment << "\n";
    f.precision( 17 );
    f << "bravais_vectors\n";
    for( int k = 0; k < 3; ++k )
        f << g.bravais[k][0] << " " << g.bravais[k][1] << " " << g.bravais[k][2] << "\n";
    f << "lattice_constant " << g.lattice_constant << "\n";
    f << "n_basis_cells " << g.n_cells[0] << " " << g.n_cells[1] << " " << g.n_cells[2] << "\n";
    f << "basis\n" << g.nb << "\n";
    for( auto & a : g.cell_atoms )
        f << a[0] << " " << a[1] << " " << a[2] << "\n";
    f << "mu_s";
    for( double m : g.mu_basis )
        f << " " << m;
    f << "\nboundary_conditions " << h.bc[0] << " " << h.bc[1] << " " << h.bc[2] << "\n";
    f << "external_field_magnitude " << h.field_magnitude / mu_B << "\n";
    f << "external_field_normal " << h.field_normal[0] << " " << h.field_normal[1] << " " << h.field_normal[2]
      << "\n";
    if( !h.aniso_K.empty() )
    {
        f << "anisotropy_magnitude " << h.aniso_K[0] << "\n";
        f << "anisotropy_normal " << h.aniso_n[0][0] << " " << h.aniso_n[0][1] << " " << h.aniso_n[0][2] << "\n";
    }
    if( !h.cubic_K.empty() )
        f << "cubic_anisotropy_magnitude " << h.cubic_K[0] << "\n";
    f << "hamiltonian heisenberg_neighbours\n";
    f << "n_shells_exchange " << h.exchange_shell_J.size() << "\njij";
    for( double v : h.exchange_shell_J )
        f << " " << v;
    f << "\nn_shells_dmi " << h.dmi_shell_D.size() << "\ndij";
    for( double v : h.dmi_shell_D )
        f << " " << v;
    f << "\ndm_chirality " << h.dmi_chirality << "\n";
    const char * dm[] = { "none", "fft", "fmm", "cutoff" };
    f << "ddi_method " << dm[std::max( 0, std::min( 3, h.ddi_method ) )] << "\n";
    f << "ddi_n_periodic_images " << h.ddi_images[0] << " " << h.ddi_images[1] << " " << h.ddi_images[2] << "\n";
    f << "ddi_pb_zero_padding " << ( h.ddi_pb_zero_padding ? 1 : 0 ) << "\n";
    auto & p = *img.llg;
    f << "llg_seed " << p.seed << "\nllg_dt " << p.dt << "\nllg_damping " << p.damping << "\nllg_temperature "
      << p.temperature << "\nllg_force_convergence " << p.force_convergence << "\nllg_n_iterations "
      << p.n_iterations << "\nllg_n_iterations_log " << p.n_iterations_log << "\n";
}
catch( ... )
{
    handle_exception( -1, -1 );
}

extern "C" const char * State_DateTime( State * state ) noexcept
try
{
    return state->datetime.c_str();
}
catch( ... )
{
    handle_exception( -1, -1 );
    return "";
}

// ===========================================================================
// Simulation.h
// ===========================================================================
extern "C" void free_run_info( Simulation_Run_Info info ) noexcept
{
    delete[] info.history_energy;
    delete[] info.history_iteration;
    delete[] info.history_max_torque;
}

static const char * solver_name( int s )
{
    switch( s )
    {
        case 0: return "VP";
        case 1: return "SIB";
        case 2: return "Depondt";
        case 3: return "Heun";
        case 4: return "RK4";
        case 5: return "LBFGS_OSO";
        case 6: return "LBFGS_Atlas";
        case 7: return "VP_OSO";
        default: return "unknown";
    }
}

// Solver::SolverFullName of the reference's solvers (core/include/engine/Solver_*.hpp)
static std::string solver_full_name( int s )
{
    switch( s )
    {
        case SOLVER_SIB: return "Semi-implicit B";
        case SOLVER_HEUN: return "Heun";
        case SOLVER_DEPONDT: return "Depondt";
        case SOLVER_RK4: return "Runge Kutta (4th order)";
        case SOLVER_VP: return "Velocity Projection";
        case SOLVER_VP_OSO: return "Velocity Projection using exponential transforms";
        case SOLVER_LBFGS_OSO: return "Limited memory Broyden-Fletcher-Goldfarb-Shanno using exponential transforms";
        case SOLVER_LBFGS_ATLAS: return "Limited memory Broyden-Fletcher-Goldfarb-Shanno using stereographic atlas";
        default: return "unknown";
    }
}

// Finish a method: copy results to the host mirror and release the device copy.
static void llg_finalize( MethodLLG & m )
{
    if( m.runner )
    {
        m.runner->download_spins();
        m.runner->download_effective_field( 0, m.image->eff_field.data() );
        m.runner.reset();
    }
    m.device_newer                = false;
    m.image->iteration_allowed    = false;
    m.image->singleshot_allowed   = false;
}

static bool llg_continue( const MethodLLG & m )
{
    return m.iteration < m.n_iterations && m.image->iteration_allowed && !stop_file_present()
           && !( m.image->max_torque < m.image->llg->force_convergence );
}

static bool walltime_expired( const MethodLLG & m )
{
    if( m.image->llg->max_walltime_sec <= 0 )
        return false;
    double s = std::chrono::duration<double>( std::chrono::steady_clock::now() - m.t_start ).count();
    return s > (double)m.image->llg->max_walltime_sec;
}

// A run that failed part-way: the image is free again and the device copy is dropped
// (its spins stay as the host mirror last held them).
static void llg_abandon( MethodLLG & m )
{
    m.runner.reset();
    m.device_newer              = false;
    m.image->iteration_allowed  = false;
    m.image->singleshot_allowed = false;
}

// Method_LLG::Save_Current (Method_LLG.cpp:310-495): the history entry, then the configuration
// and energy files the output parameters ask for.  The host mirror must hold the spins.
static void llg_save_current( MethodLLG & m, long iteration, bool initial, bool final )
{
    m.history_iteration.push_back( (int)iteration );
    m.history_max_torque.push_back( (float)m.image->max_torque );
    m.history_energy.push_back( (float)m.image->E );
    if( !m.image->llg->output_any )
        return;
    llg_save_configuration( *m.image, m.idx_image, m.starttime, iteration, m.n_iterations, m.image->max_torque,
                            solver_full_name( m.solver ), initial, final );
    llg_save_energy( *m.image, m.idx_image, m.starttime, iteration, m.n_iterations, initial, final );
}

// Whether the step files of the output parameters are written at log steps.
static bool llg_step_files( const LLGParameters & P )
{
    return P.output_any
           && ( P.output_configuration_step || P.output_configuration_archive || P.output_energy_step
                || P.output_energy_archive );
}

// The iteration loop of Method::Iterate (Method.cpp:71-102) saves after the body that starts
// at iteration `it` when it > 0 and it is a multiple of n_iterations_log.  Of the next
// `chunks` bodies of `am` steps from `iteration` on, the number up to and including the first
// such body (chunks if there is none).
static long chunks_to_log_body( long iteration, long chunks, long am, long log )
{
    for( long k = 1; k <= chunks; ++k )
    {
        const long it = iteration + ( k - 1 ) * am;
        if( it > 0 && it % log == 0 )
            return k;
    }
    return chunks;
}

static void fill_run_info( Simulation_Run_Info * info, const MethodLLG & m )
{
    if( !info )
        return;
    info->max_torque       = (float)m.image->max_torque;
    info->total_iterations = (int)m.iteration;
    double ms = std::chrono::duration<double, std::milli>( std::chrono::steady_clock::now() - m.t_start ).count();
    info->total_walltime = (int)ms;
    info->total_ips      = ms > 0 ? float( m.iteration / ms * 1000.0 ) : 0.f;
    // Simulation.cpp:49-70: arrays owned by the caller, released by free_run_info
    const int n = (int)m.history_iteration.size();
    if( n > 0 )
    {
        info->n_history_iteration  = n;
        info->history_iteration    = new int[n];
        info->n_history_max_torque = n;
        info->history_max_torque   = new float[n];
        info->n_history_energy     = n;
        info->history_energy       = new float[n];
        std::copy( m.history_iteration.begin(), m.history_iteration.end(), info->history_iteration );
        std::copy( m.history_max_torque.begin(), m.history_max_torque.end(), info->history_max_torque );
        std::copy( m.history_energy.begin(), m.history_energy.end(), info->history_energy );
    }
}

extern "C" void Simulation_LLG_Start( State * state, int solver_type, int n_iterations, int n_iterations_log,
                                      bool singleshot, Simulation_Run_Info * info, int idx_image,
                                      int idx_chain ) noexcept
try
{
    std::shared_ptr<Image> image;
    std::shared_ptr<Chain> chain;
    from_indices( state, idx_image, idx_chain, image, chain );
    if( image->iteration_allowed || chain->iteration_allowed )
        throw api_error( 3, "Tried to use Simulation_Start on image " + std::to_string( idx_image ) + " of chain "
                                + std::to_string( idx_chain ) + ", but there is already a simulation running." );
    auto m = std::make_shared<MethodLLG>();
    {
        std::lock_guard<std::recursive_mutex> lk( image->mtx );
        image->iteration_allowed  = true;
        image->singleshot_allowed = singleshot;
        if( n_iterations > 0 )
            image->llg->n_iterations = n_iterations;
        if( n_iterations_log > 0 )
            image->llg->n_iterations_log = n_iterations_log;
        m->image            = image;
        m->solver           = solver_type;
        m->solver_name      = solver_name( solver_type );
        m->n_iterations     = std::max( 1L, image->llg->n_iterations );
        m->n_iterations_log = std::min( image->llg->n_iterations_log, m->n_iterations );
        if( m->n_iterations_log <= 0 )
            m->n_iterations_log = m->n_iterations;
        m->amortize = std::max( 1L, std::min( image->llg->n_iterations_amortize, m->n_iterations_log ) );
        try
        {
            ImageSpec sp{ image->geometry, image->hamiltonian, *image->llg, image->spins.data() };
            m->runner.reset( new LLGRunner( { sp }, solver_type ) );
        }
        catch( ... )
        {
            image->iteration_allowed  = false;
            image->singleshot_allowed = false;
            throw;
        }
        // constructor force evaluation and Hook_Post_Iteration (Method_LLG.cpp:22-63)
        m->runner->initial_observe();
        image->E          = m->runner->energy( 0 );
        image->max_torque = m->runner->max_torque( 0 );
        m->picoseconds    = image->llg->dt;
        m->t_start        = std::chrono::steady_clock::now();
    }
    state->method_image[idx_image] = m;
    m->idx_image = idx_image < 0 ? state->idx_active_image : idx_image;
    m->starttime = current_date_time();
    m->iteration = 0;
    try
    {
        // initial save, also in single-shot mode (Simulation.cpp:24-38)
        llg_save_current( *m, 0, true, false );
        if( singleshot )
            return;

        // Method::Iterate, Method.cpp:57-113: n_iterations_amortize steps per body, then
        // Hook_Post_Iteration (energy, max torque and the convergence test, evaluated on the
        // device after every body).  The host submits batches of bodies and checks the stop
        // requests (Simulation_Stop, the STOP file, the wall time) between them; a batch lasts
        // about 0.1 s.  With step files, a batch ends on the next log body and the spins come back.
        const long am       = m->amortize;
        const long log      = m->n_iterations_log;
        const bool step_out = llg_step_files( *image->llg );
        long chunks         = std::max( 1L, 16 / am );
        while( llg_continue( *m ) && !walltime_expired( *m ) )
        {
            const long left = ( m->n_iterations - m->iteration + am - 1 ) / am;
            long outer      = std::min( left, chunks );
            if( step_out )
                outer = chunks_to_log_body( m->iteration, outer, am, log );
            auto t0    = std::chrono::steady_clock::now();
            long done  = m->runner->run( outer * am, (int)am, true, image->llg->force_convergence );
            auto t1    = std::chrono::steady_clock::now();
            double sec = std::chrono::duration<double>( t1 - t0 ).count();
            if( sec > 0 && done > 0 )
            {
                m->ips          = done / sec;
                const double pb = sec / (double)( ( done + am - 1 ) / am ); // seconds per body
                chunks          = std::max( 1L, std::min( 1L << 20, (long)( 0.1 / pb ) ) );
            }
            m->iteration += done;
            m->picoseconds += done * image->llg->dt;
            image->E          = m->runner->energy( 0 );
            image->max_torque = m->runner->max_torque( 0 );
            const long it     = m->iteration - am; // first iteration of the last body
            if( step_out && done > 0 && it > 0 && it % log == 0 )
            {
                m->runner->download_spins();
                llg_save_current( *m, it, false, false );
            }
        }
        llg_finalize( *m );
    }
    catch( ... )
    {
        llg_abandon( *m );
        throw;
    }
    llg_save_current( *m, m->iteration, false, true );
    fill_run_info( info, *m );
}
catch( ... )
{
    handle_exception( idx_image, idx_chain );
}

// Independent LLG runs of the active images of several states, batched into one runner: every
// stage of every image is one launch, each image keeps its own Hamiltonian, parameters,
// convergence test and iteration count (as if Simulation_LLG_Start ran on each in turn).
extern "C" void Spirit_HIP_Simulation_LLG_Start_Batch( State ** states, int n_states, int solver_type,
                                                      int n_iterations, int n_iterations_log,
                                                      Simulation_Run_Info * info ) noexcept
try
{
    if( !states || n_states <= 0 )
        throw api_error( 2, "Spirit_HIP_Simulation_LLG_Start_Batch: no states" );
    struct Entry
    {
        std::shared_ptr<Image> image;
        long n_iterations = 0, iteration = 0;
        bool active       = true;
    };
    std::vector<Entry> E( n_states );
    for( int k = 0; k < n_states; ++k )
    {
        int idx_image = -1, idx_chain = -1;
        std::shared_ptr<Chain> chain;
        from_indices( states[k], idx_image, idx_chain, E[k].image, chain );
        if( E[k].image->iteration_allowed || chain->iteration_allowed )
            throw api_error( 3, "Spirit_HIP_Simulation_LLG_Start_Batch: state " + std::to_string( k )
                                    + " already runs a simulation" );
        for( int j = 0; j < k; ++j )
            if( E[j].image == E[k].image )
                throw api_error( 3, "Spirit_HIP_Simulation_LLG_Start_Batch: an image appears twice" );
    }
    std::vector<ImageSpec> specs;
    long am = 1;
    for( int k = 0; k < n_states; ++k )
    {
        auto & img = *E[k].image;
        std::lock_guard<std::recursive_mutex> lk( img.mtx );
        img.iteration_allowed  = true;
        img.singleshot_allowed = false;
        if( n_iterations > 0 )
            img.llg->n_iterations = n_iterations;
        if( n_iterations_log > 0 )
            img.llg->n_iterations_log = n_iterations_log;
        E[k].n_iterations = std::max( 1L, img.llg->n_iterations );
        if( k == 0 )
            am = std::max( 1L, std::min( img.llg->n_iterations_amortize, E[k].n_iterations ) );
        specs.push_back( ImageSpec{ img.geometry, img.hamiltonian, *img.llg, img.spins.data() } );
    }
    std::unique_ptr<LLGRunner> runner;
    auto release = [&]() {
        for( auto & e : E )
        {
            e.image->iteration_allowed  = false;
            e.image->singleshot_allowed = false;
        }
    };
    try
    {
        runner.reset( new LLGRunner( specs, solver_type ) );
        runner->initial_observe();
    }
    catch( ... )
    {
        release();
        throw;
    }
    auto t_start = std::chrono::steady_clock::now();
    for( int k = 0; k < n_states; ++k )
    {
        E[k].image->E          = runner->energy( k );
        E[k].image->max_torque = runner->max_torque( k );
    }
    // Method::Iterate per image; convergence is tested on the device every am steps.  Batches
    // last about 0.1 s, so stop requests are seen at least that often.
    long chunks = std::max( 1L, 16 / am );
    try
    {
        while( true )
        {
            long batch = 0;
            for( int k = 0; k < n_states; ++k )
            {
                auto & img = *E[k].image;
                const bool walltime
                    = img.llg->max_walltime_sec > 0
                      && std::chrono::duration<double>( std::chrono::steady_clock::now() - t_start ).count()
                             > (double)img.llg->max_walltime_sec;
                E[k].active = E[k].active && E[k].iteration < E[k].n_iterations && img.iteration_allowed
                              && !stop_file_present() && !( img.max_torque < img.llg->force_convergence ) && !walltime;
                runner->set_image_stopped( k, !E[k].active );
                if( E[k].active )
                {
                    long left = ( E[k].n_iterations - E[k].iteration + am - 1 ) / am * am;
                    batch     = batch == 0 ? left : std::min( batch, left );
                }
            }
            if( batch == 0 )
                break;
            batch       = std::min( batch, chunks * am );
            auto t0     = std::chrono::steady_clock::now();
            long done   = runner->run( batch, (int)am, true, 0.0 );
            double sec  = std::chrono::duration<double>( std::chrono::steady_clock::now() - t0 ).count();
            if( sec > 0 && done > 0 )
                chunks = std::max( 1L, std::min( 1L << 20, (long)( 0.1 / ( sec / (double)( ( done + am - 1 ) / am ) ) ) ) );
            for( int k = 0; k < n_states; ++k )
            {
                E[k].iteration += runner->steps_last_run( k );
                E[k].image->E          = runner->energy( k );
                E[k].image->max_torque = runner->max_torque( k );
            }
        }
        runner->download_spins();
        std::vector<double *> eff( n_states );
        for( int k = 0; k < n_states; ++k )
            eff[k] = E[k].image->eff_field.data();
        runner->download_effective_fields( eff );
    }
    catch( ... )
    {
        release();
        throw;
    }
    release();
    double ms = std::chrono::duration<double, std::milli>( std::chrono::steady_clock::now() - t_start ).count();
    if( info )
        for( int k = 0; k < n_states; ++k )
        {
            info[k].max_torque       = (float)E[k].image->max_torque;
            info[k].total_iterations = (int)E[k].iteration;
            info[k].total_walltime   = (int)ms;
            info[k].total_ips        = ms > 0 ? float( E[k].iteration / ms * 1000.0 ) : 0.f;
        }
}
catch( ... )
{
    handle_exception( -1, -1 );
}

// Colour classes for the parallel Metropolis sweep: colour(a, b, c) = (ca a + cb b + cc c) mod K
// with no exchange or DMI pair inside a class (and consistent across periodic wraps).  One
// basis atom; the smallest K <= 8 that works.
static bool mc_colouring( const Geometry & g, const Hamiltonian & h, int & K, int c[3] )
{
    if( g.nb != 1 )
        return false;
    std::vector<std::array<int, 3>> ts;
    auto add = [&]( const std::vector<Pair> & v ) {
        for( auto & p : v )
        {
            bool connects = false, self = true;
            for( int d = 0; d < 3; ++d )
            {
                if( p.t[d] != 0 && g.n_cells[d] > 1 && std::abs( p.t[d] ) < g.n_cells[d] )
                    connects = true;
                if( p.t[d] != 0 && g.n_cells[d] > 1 )
                    self = false;
            }
            if( connects && !self )
                ts.push_back( { p.t[0], p.t[1], p.t[2] } );
        }
    };
    add( h.exch_pairs );
    add( h.dmi_pairs );
    for( K = 2; K <= 8; ++K )
        for( c[0] = 0; c[0] < K; ++c[0] )
            for( c[1] = 0; c[1] < K; ++c[1] )
                for( c[2] = 0; c[2] < K; ++c[2] )
                {
                    bool ok = true;
                    for( int d = 0; d < 3 && ok; ++d )
                        if( h.bc[d] && g.n_cells[d] > 1 && ( c[d] * g.n_cells[d] ) % K != 0 )
                            ok = false;
                    for( auto & t : ts )
                    {
                        if( !ok )
                            break;
                        const int v = ( ( c[0] * t[0] + c[1] * t[1] + c[2] * t[2] ) % K + K ) % K;
                        if( v == 0 )
                            ok = false;
                    }
                    if( ok )
                        return true;
                }
    return false;
}

// Method_MC's set-up (Method_MC.cpp:20-46) on the GPU: the colour classes of the lattice, the
// image's device record and spins, the starting cone angle and the Metropolis kernels' arguments.
static std::shared_ptr<MethodMC> mc_setup( State * state, const std::shared_ptr<Image> & image, int idx_image )
{
    MCParameters & P = *image->mc;
    int K = 0, col[3] = { 0, 0, 0 };
    if( !mc_colouring( *image->geometry, *image->hamiltonian, K, col ) )
        throw api_error( 2, "Simulation_MC_Start: the lattice's pairs admit no colouring into at most 8 classes "
                            "(one basis atom is required)" );
    sync_image_host( state, idx_image );
    auto m   = std::make_shared<MethodMC>();
    m->image = image;
    ImageSpec sp{ image->geometry, image->hamiltonian, *image->llg, image->spins.data() };
    m->runner.reset( new LLGRunner( { sp }, SOLVER_DEPONDT ) );
    const int nos      = image->nos();
    const int n_blocks = ( nos + 255 ) / 256;
    m->cone.resize( 2 * sizeof( double ) );
    m->part.resize( sizeof( double ) * K * (size_t)n_blocks );
    m->step.resize( sizeof( uint64_t ) );
    m->cone0 = Pi * P.metropolis_cone_angle / 180.0; // Method_MC.cpp:34
    HIPCHECK( hipMemcpy( m->cone.as<void>(), &m->cone0, sizeof( double ), hipMemcpyHostToDevice ) );
    HIPCHECK( hipMemset( m->part.as<void>(), 0, m->part.size() ) );
    HIPCHECK( hipMemset( m->step.as<void>(), 0, sizeof( uint64_t ) ) );
    McArgs & A    = m->A;
    A.img         = m->runner->dev_images();
    A.s           = m->runner->dev_spins();
    A.nos         = nos;
    A.K           = K;
    A.ca          = col[0];
    A.cb          = col[1];
    A.cc          = col[2];
    A.step_cone   = P.metropolis_step_cone ? 1 : 0;
    A.adaptive    = P.metropolis_cone_adaptive ? 1 : 0;
    A.seed        = P.seed;
    A.temperature = P.temperature;
    A.kT          = k_B * P.temperature;
    A.target      = P.acceptance_ratio_target;
    A.cone        = m->cone.as<double>();
    A.part        = m->part.as<double>();
    A.n_blocks    = n_blocks;
    A.step        = m->step.as<uint64_t>();
    // per-colour sweeps where ca is invertible mod K (one basis atom is required above)
    for( int x = 1; x < K; ++x )
        if( ( col[0] * x ) % K == 1 )
            A.inv_ca = x;
    A.groups        = ( image->geometry->n_cells[0] + K - 1 ) / K;
    A.n_blocks_rows = (int)( ( (long)A.groups * image->geometry->n_cells[1] * image->geometry->n_cells[2] + 255 )
                             / 256 );
    if( A.n_blocks_rows > n_blocks )
        A.inv_ca = 0;
    m->n_iterations = std::max( 1L, P.n_iterations );
    m->t_start      = std::chrono::steady_clock::now();
    return m;
}

// n Metropolis iterations (Method_MC::Iteration, Method_MC.cpp:40-183) on the device
static void mc_sweeps( MethodMC & m, long n )
{
    for( long k = 0; k < n; ++k )
        mc_launch_iteration( m.A, m.step.as<uint64_t>(), m.runner->stream );
    HIPCHECK( hipGetLastError() );
    HIPCHECK( hipStreamSynchronize( m.runner->stream ) );
    m.iteration += n;
}

// The spins and, with an adaptive cone, the cone angle back to the host; the image's energy.
static void mc_download( MethodMC & m )
{
    m.runner->download_spins();
    MCParameters & P = *m.image->mc;
    double cone_end  = m.cone0;
    HIPCHECK( hipMemcpy( &cone_end, m.cone.as<void>(), sizeof( double ), hipMemcpyDeviceToHost ) );
    if( P.metropolis_step_cone && P.metropolis_cone_adaptive )
        P.metropolis_cone_angle = cone_end * 180.0 / Pi;
    m.image->E = update_image_energy( *m.image, true );
}

// Method_MC::Finalize (Method_MC.cpp:203-206) of a single-shot run
static void mc_finalize( State * state, int idx_image )
{
    auto it = state->method_mc.find( idx_image );
    if( it == state->method_mc.end() )
        return;
    it->second->image->iteration_allowed  = false;
    it->second->image->singleshot_allowed = false;
    state->method_mc.erase( it );
}

// Method_MC with the Metropolis algorithm (Method_MC.cpp:40-200, Method.cpp:57-113): n_iterations
// sweeps on the GPU; the spins, the adapted cone angle and the energy come back at the end.  In
// single-shot mode the method is set up and Simulation_SingleShot / _N_Shot run the sweeps.
extern "C" void Simulation_MC_Start( State * state, int n_iterations, int n_iterations_log, bool singleshot,
                                     Simulation_Run_Info * info, int idx_image, int idx_chain ) noexcept
try
{
    std::shared_ptr<Image> image;
    std::shared_ptr<Chain> chain;
    from_indices( state, idx_image, idx_chain, image, chain );
    if( image->iteration_allowed || chain->iteration_allowed )
        throw api_error( 3, "Tried to use Simulation_Start on image " + std::to_string( idx_image ) + " of chain "
                                + std::to_string( idx_chain ) + ", but there is already a simulation running." );
    (void)n_iterations_log;
    MCParameters & P = *image->mc;
    if( n_iterations > 0 )
        P.n_iterations = n_iterations;
    std::lock_guard<std::recursive_mutex> lk( image->mtx );
    auto m = mc_setup( state, image, idx_image );
    if( singleshot )
    {
        state->method_mc[idx_image] = m;
        image->singleshot_allowed   = true;
        image->iteration_allowed    = true;
        return;
    }
    image->iteration_allowed = true;
    const long n             = m->n_iterations;
    try
    {
        const long batch = 1000;
        while( m->iteration < n && image->iteration_allowed && !stop_file_present() )
        {
            mc_sweeps( *m, std::min( batch, n - m->iteration ) );
            if( P.max_walltime_sec > 0
                && std::chrono::duration<double>( std::chrono::steady_clock::now() - m->t_start ).count()
                       > (double)P.max_walltime_sec )
                break;
        }
        mc_download( *m );
    }
    catch( ... )
    {
        image->iteration_allowed = false;
        throw;
    }
    image->iteration_allowed = false;
    if( info )
    {
        double ms = std::chrono::duration<double, std::milli>( std::chrono::steady_clock::now() - m->t_start ).count();
        info->total_iterations = (int)m->iteration;
        info->total_walltime   = (int)ms;
        info->total_ips        = ms > 0 ? float( m->iteration / ms * 1000.0 ) : 0.f;
    }
}
catch( ... )
{
    handle_exception( idx_image, idx_chain );
}

// Simulation_N_Shot of a single-shot Monte Carlo run (Simulation.cpp:454-535): N iterations
// while the method continues, the spins back after them, Finalize once it stops.
static void mc_n_shot( State * state, int N, int idx_image )
{
    auto m           = state->method_mc[idx_image];
    Image & image    = *m->image;
    MCParameters & P = *image.mc;
    auto walltime    = [&]() {
        return P.max_walltime_sec > 0
               && std::chrono::duration<double>( std::chrono::steady_clock::now() - m->t_start ).count()
                      > (double)P.max_walltime_sec;
    };
    auto cont = [&]() { return m->iteration < m->n_iterations && image.iteration_allowed && !stop_file_present(); };
    try
    {
        if( cont() && !walltime() )
        {
            std::lock_guard<std::recursive_mutex> lk( image.mtx );
            // all N iterations, as the reference's loop (the continuation is tested before it)
            const long n = std::max( 0, N );
            auto t0      = std::chrono::steady_clock::now();
            mc_sweeps( *m, n );
            const double sec = std::chrono::duration<double>( std::chrono::steady_clock::now() - t0 ).count();
            if( sec > 0 && n > 0 )
                m->ips = n / sec;
            mc_download( *m );
        }
        if( !cont() || walltime() )
            mc_finalize( state, idx_image );
    }
    catch( ... )
    {
        mc_finalize( state, idx_image );
        throw;
    }
}

extern "C" void Simulation_MMF_Start( State * state, int solver_type, int n_iterations, int n_iterations_log,
                                      bool singleshot, Simulation_Run_Info * info, int idx_image,
                                      int idx_chain ) noexcept
{
    log_message( 2, "Simulation_MMF_Start: minimum mode following is not part of the HIP engine" );
}

extern "C" void Simulation_EMA_Start( State * state, int n_iterations, int n_iterations_log, bool singleshot,
                                      Simulation_Run_Info * info, int idx_image, int idx_chain ) noexcept
{
    log_message( 2, "Simulation_EMA_Start: eigenmode analysis is not part of the HIP engine" );
}

extern "C" void Simulation_N_Shot( State * state, int N, int idx_image, int idx_chain ) noexcept;

extern "C" void Simulation_SingleShot( State * state, int idx_image, int idx_chain ) noexcept
{
    Simulation_N_Shot( state, 1, idx_image, idx_chain );
}

// Simulation.cpp:454-540
extern "C" void Simulation_N_Shot( State * state, int N, int idx_image, int idx_chain ) noexcept
try
{
    std::shared_ptr<Image> image;
    std::shared_ptr<Chain> chain;
    from_indices( state, idx_image, idx_chain, image, chain );
    if( chain->iteration_allowed && chain->singleshot_allowed )
    {
        gneb_n_shot( state, N );
        return;
    }
    if( !( image->iteration_allowed && image->singleshot_allowed ) )
        throw api_error( 3, "Tried to use Simulation_SingleShot on image " + std::to_string( idx_image )
                                + " but no SingleShot simulation has been started." );
    if( state->method_mc.count( idx_image ) )
    {
        mc_n_shot( state, N, idx_image );
        return;
    }
    auto m = state->method_image[idx_image];
    try
    {
        if( llg_continue( *m ) && !walltime_expired( *m ) )
        {
            // Simulation.cpp:477-504: N iterations, each followed by Hook_Post_Iteration (energy,
            // max torque and effective field of every step; convergence is only tested after the
            // N steps), saving after the iteration it when it > 0 is a multiple of n_iterations_log
            std::lock_guard<std::recursive_mutex> lk( image->mtx );
            const bool step_out = llg_step_files( *image->llg );
            long left           = std::max( 0, N );
            while( left > 0 )
            {
                const long n = step_out ? chunks_to_log_body( m->iteration, left, 1, m->n_iterations_log ) : left;
                auto t0      = std::chrono::steady_clock::now();
                m->runner->run( n, 1, false, image->llg->force_convergence );
                double sec = std::chrono::duration<double>( std::chrono::steady_clock::now() - t0 ).count();
                if( sec > 0 )
                    m->ips = n / sec;
                m->iteration += n;
                m->picoseconds += n * image->llg->dt;
                image->E          = m->runner->energy( 0 );
                image->max_torque = m->runner->max_torque( 0 );
                m->device_newer   = true;
                left -= n;
                const long it = m->iteration - 1;
                if( step_out && it > 0 && it % m->n_iterations_log == 0 )
                {
                    sync_image_host( state, idx_image );
                    llg_save_current( *m, it, false, false );
                }
            }
            if( state->host_sync )
                sync_image_host( state, idx_image );
        }
        if( !llg_continue( *m ) || walltime_expired( *m ) )
        {
            llg_finalize( *m );
            llg_save_current( *m, m->iteration, false, true );
        }
    }
    catch( ... )
    {
        llg_abandon( *m );
        throw;
    }
}
catch( ... )
{
    handle_exception( idx_image, idx_chain );
}

extern "C" void Simulation_Stop( State * state, int idx_image, int idx_chain ) noexcept
try
{
    std::shared_ptr<Image> image;
    std::shared_ptr<Chain> chain;
    from_indices( state, idx_image, idx_chain, image, chain );
    if( image->iteration_allowed )
    {
        std::lock_guard<std::recursive_mutex> lk( image->mtx );
        image->iteration_allowed = false;
        if( image->singleshot_allowed )
        {
            if( state->method_mc.count( idx_image ) )
                mc_finalize( state, idx_image );
            else
            {
                auto m = state->method_image[idx_image];
                if( m )
                    llg_finalize( *m );
            }
        }
    }
    else if( chain->iteration_allowed )
    {
        chain->iteration_allowed = false;
        if( chain->singleshot_allowed )
            gneb_finalize( state );
    }
}
catch( ... )
{
    handle_exception( idx_image, idx_chain );
}

extern "C" void Simulation_Stop_All( State * state ) noexcept
try
{
    for( int i = 0; i < state->chain->noi(); ++i )
        Simulation_Stop( state, i, -1 );
    Simulation_Stop( state, -1, -1 );
}
catch( ... )
{
    handle_exception( -1, -1 );
}

extern "C" float Simulation_Get_MaxTorqueComponent( State * state, int idx_image, int idx_chain ) noexcept
try
{
    std::shared_ptr<Image> image;
    std::shared_ptr<Chain> chain;
    from_indices( state, idx_image, idx_chain, image, chain );
    if( chain->iteration_allowed && state->method_chain )
        return (float)state->method_chain->max_torque;
    return (float)image->max_torque;
}
catch( ... )
{
    handle_exception( idx_image, idx_chain );
    return 0;
}

extern "C" float Simulation_Get_MaxTorqueNorm( State * state, int idx_image, int idx_chain ) noexcept
{
    return Simulation_Get_MaxTorqueComponent( state, idx_image, idx_chain );
}

extern "C" void Simulation_Get_Chain_MaxTorqueComponents( State * state, float * torques, int idx_chain ) noexcept
try
{
    if( state->method_chain )
        for( size_t i = 0; i < state->method_chain->max_torque_all.size(); ++i )
            torques[i] = (float)state->method_chain->max_torque_all[i];
}
catch( ... )
{
    handle_exception( -1, idx_chain );
}

extern "C" void Simulation_Get_Chain_MaxTorqueNorms( State * state, float * torques, int idx_chain ) noexcept
{
    Simulation_Get_Chain_MaxTorqueComponents( state, torques, idx_chain );
}

extern "C" float Simulation_Get_IterationsPerSecond( State * state, int idx_image, int idx_chain ) noexcept
try
{
    std::shared_ptr<Image> image;
    std::shared_ptr<Chain> chain;
    from_indices( state, idx_image, idx_chain, image, chain );
    if( chain->iteration_allowed && state->method_chain )
        return (float)state->method_chain->ips;
    auto mc = state->method_mc.find( idx_image );
    if( mc != state->method_mc.end() )
        return (float)mc->second->ips;
    auto it = state->method_image.find( idx_image );
    if( it != state->method_image.end() && it->second )
        return (float)it->second->ips;
    return 0;
}
catch( ... )
{
    handle_exception( idx_image, idx_chain );
    return 0;
}

extern "C" int Simulation_Get_Iteration( State * state, int idx_image, int idx_chain ) noexcept
try
{
    std::shared_ptr<Image> image;
    std::shared_ptr<Chain> chain;
    from_indices( state, idx_image, idx_chain, image, chain );
    auto mc = state->method_mc.find( idx_image );
    if( mc != state->method_mc.end() )
        return (int)mc->second->iteration;
    if( state->method_chain && ( chain->iteration_allowed || !state->method_image.count( idx_image ) ) )
        return (int)state->method_chain->iteration;
    auto it = state->method_image.find( idx_image );
    if( it != state->method_image.end() && it->second )
        return (int)it->second->iteration;
    return 0;
}
catch( ... )
{
    handle_exception( idx_image, idx_chain );
    return 0;
}

extern "C" float Simulation_Get_Time( State * state, int idx_image, int idx_chain ) noexcept
try
{
    std::shared_ptr<Image> image;
    std::shared_ptr<Chain> chain;
    from_indices( state, idx_image, idx_chain, image, chain );
    auto it = state->method_image.find( idx_image );
    if( it != state->method_image.end() && it->second )
        return (float)it->second->picoseconds;
    return 0;
}
catch( ... )
{
    handle_exception( idx_image, idx_chain );
    return 0;
}

extern "C" int Simulation_Get_Wall_Time( State * state, int idx_image, int idx_chain ) noexcept
try
{
    std::shared_ptr<Image> image;
    std::shared_ptr<Chain> chain;
    from_indices( state, idx_image, idx_chain, image, chain );
    auto mc = state->method_mc.find( idx_image );
    if( mc != state->method_mc.end() )
        return (int)std::chrono::duration<double, std::milli>( std::chrono::steady_clock::now()
                                                               - mc->second->t_start )
            .count();
    auto it = state->method_image.find( idx_image );
    if( it != state->method_image.end() && it->second )
        return (int)std::chrono::duration<double, std::milli>( std::chrono::steady_clock::now()
                                                               - it->second->t_start )
            .count();
    return 0;
}
catch( ... )
{
    handle_exception( idx_image, idx_chain );
    return 0;
}

extern "C" const char * Simulation_Get_Solver_Name( State * state, int idx_image, int idx_chain ) noexcept
try
{
    std::shared_ptr<Image> image;
    std::shared_ptr<Chain> chain;
    from_indices( state, idx_image, idx_chain, image, chain );
    if( chain->iteration_allowed && state->method_chain )
        return state->method_chain->solver_name.c_str();
    auto it = state->method_image.find( idx_image );
    if( it != state->method_image.end() && it->second && image->iteration_allowed )
        return it->second->solver_name.c_str();
    return "";
}
catch( ... )
{
    handle_exception( idx_image, idx_chain );
    return "";
}

extern "C" const char * Simulation_Get_Method_Name( State * state, int idx_image, int idx_chain ) noexcept
try
{
    std::shared_ptr<Image> image;
    std::shared_ptr<Chain> chain;
    from_indices( state, idx_image, idx_chain, image, chain );
    if( chain->iteration_allowed )
        return "GNEB";
    if( image->iteration_allowed )
        return state->method_mc.count( idx_image ) ? "MC" : "LLG";
    return "None";
}
catch( ... )
{
    handle_exception( idx_image, idx_chain );
    return "";
}

extern "C" bool Simulation_Running_On_Image( State * state, int idx_image, int idx_chain ) noexcept
try
{
    std::shared_ptr<Image> image;
    std::shared_ptr<Chain> chain;
    from_indices( state, idx_image, idx_chain, image, chain );
    return image->iteration_allowed;
}
catch( ... )
{
    handle_exception( idx_image, idx_chain );
    return false;
}

extern "C" bool Simulation_Running_On_Chain( State * state, int idx_chain ) noexcept
try
{
    return state->chain->iteration_allowed;
}
catch( ... )
{
    handle_exception( -1, idx_chain );
    return false;
}

extern "C" bool Simulation_Running_Anywhere_On_Chain( State * state, int idx_chain ) noexcept
try
{
    if( state->chain->iteration_allowed )
        return true;
    for( auto & im : state->chain->images )
        if( im->iteration_allowed )
            return true;
    return false;
}
catch( ... )
{
    handle_exception( -1, idx_chain );
    return false;
}
