// C API: the GNEB method on the chain (Simulation_GNEB_Start) and chain data updates.
#include "api_common.hpp"

using namespace spirit_hip;

namespace spirit_hip
{

void gneb_n_shot( State * state, int N )
{
    throw api_error( 2, "GNEB single-shot: no GNEB method is running" );
}

void gneb_finalize( State * state )
{
    state->chain->iteration_allowed  = false;
    state->chain->singleshot_allowed = false;
}

void gneb_sync_host( State * state ) {}

} // namespace spirit_hip

extern "C" void Simulation_GNEB_Start( State * state, int solver_type, int n_iterations, int n_iterations_log,
                                       bool singleshot, Simulation_Run_Info * info, int idx_chain ) noexcept
{
    log_message( 2, "Simulation_GNEB_Start: GNEB is not available yet in this build" );
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
