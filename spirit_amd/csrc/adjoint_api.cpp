// C API (engine extension, include/Spirit/Spirit_HIP.h): differentiable LLG runs of one image.
#include "adjoint.hpp"
#include "api_common.hpp"

using namespace spirit_hip;

struct Spirit_HIP_Adjoint
{
    std::unique_ptr<AdjointRunner> r;
};

namespace
{

// AoS spins (host or device) -> device SoA
void spins_in( AdjointRunner & r, const double * src, DeviceBuffer & dst, int on_device )
{
    const long long n = r.nos();
    dst.resize( sizeof( double ) * 3 * (size_t)n );
    if( on_device )
        device_aos_to_soa( src, dst.as<double>(), n, r.stream() );
    else
    {
        std::vector<double> soa( 3 * (size_t)n );
        aos_to_soa( src, soa.data(), n );
        HIPCHECK( hipMemcpyAsync( dst.as<void>(), soa.data(), sizeof( double ) * soa.size(), hipMemcpyHostToDevice,
                                  r.stream() ) );
    }
    HIPCHECK( hipStreamSynchronize( r.stream() ) );
}

// device SoA -> AoS spins (host or device)
void spins_out( AdjointRunner & r, const DeviceBuffer & src, double * dst, int on_device )
{
    const long long n = r.nos();
    if( on_device )
        device_soa_to_aos( src.as<double>(), dst, n, r.stream() );
    else
    {
        std::vector<double> soa( 3 * (size_t)n );
        HIPCHECK( hipMemcpyAsync( soa.data(), src.as<void>(), sizeof( double ) * soa.size(), hipMemcpyDeviceToHost,
                                  r.stream() ) );
        HIPCHECK( hipStreamSynchronize( r.stream() ) );
        soa_to_aos( soa.data(), dst, n );
    }
    HIPCHECK( hipStreamSynchronize( r.stream() ) );
}

AdjointRunner & runner( Spirit_HIP_Adjoint * h )
{
    if( !h || !h->r )
        throw api_error( 2, "Spirit_HIP_Adjoint: invalid handle" );
    return *h->r;
}

} // namespace

extern "C" Spirit_HIP_Adjoint * Spirit_HIP_Adjoint_Create( State * state, int solver_type, int n_steps,
                                                           int checkpoint_every, int idx_image, int idx_chain ) noexcept
try
{
    std::shared_ptr<Image> image;
    std::shared_ptr<Chain> chain;
    from_indices( state, idx_image, idx_chain, image, chain );
    sync_image_host( state, idx_image );
    auto h = new Spirit_HIP_Adjoint;
    try
    {
        h->r.reset( new AdjointRunner( *image->geometry, *image->hamiltonian, *image->llg, solver_type, n_steps,
                                       checkpoint_every ) );
    }
    catch( ... )
    {
        delete h;
        throw;
    }
    return h;
}
catch( ... )
{
    handle_exception( idx_image, idx_chain );
    return nullptr;
}

extern "C" void Spirit_HIP_Adjoint_Destroy( Spirit_HIP_Adjoint * handle ) noexcept
{
    delete handle;
}

extern "C" int Spirit_HIP_Adjoint_Get_Sizes( Spirit_HIP_Adjoint * handle, int * nos, int * n_exchange, int * n_dmi,
                                             int * n_aniso ) noexcept
try
{
    auto & r = runner( handle );
    if( nos )
        *nos = r.nos();
    if( n_exchange )
        *n_exchange = r.n_exchange_shells();
    if( n_dmi )
        *n_dmi = r.n_dmi_shells();
    if( n_aniso )
        *n_aniso = r.n_aniso();
    return 0;
}
catch( ... )
{
    handle_exception( -1, -1 );
    return 1;
}

extern "C" int Spirit_HIP_Adjoint_Set_Parameters( Spirit_HIP_Adjoint * handle, const double * J, const double * D,
                                                  const double * B, const double * K, const double * K_site ) noexcept
try
{
    runner( handle ).set_parameters( J, D, B, K, K_site );
    return 0;
}
catch( ... )
{
    handle_exception( -1, -1 );
    return 1;
}

extern "C" int Spirit_HIP_Adjoint_Get_Parameters( Spirit_HIP_Adjoint * handle, double * J, double * D, double * B,
                                                  double * K, double * K_site ) noexcept
try
{
    runner( handle ).get_parameters( J, D, B, K, K_site );
    return 0;
}
catch( ... )
{
    handle_exception( -1, -1 );
    return 1;
}

extern "C" int Spirit_HIP_Adjoint_Forward( Spirit_HIP_Adjoint * handle, const double * s0, double * sN,
                                           int on_device ) noexcept
try
{
    auto & r = runner( handle );
    DeviceBuffer a, b;
    spins_in( r, s0, a, on_device );
    b.resize( a.size() );
    r.forward( a.as<double>(), b.as<double>() );
    spins_out( r, b, sN, on_device );
    return 0;
}
catch( ... )
{
    handle_exception( -1, -1 );
    return 1;
}

extern "C" int Spirit_HIP_Adjoint_Backward( Spirit_HIP_Adjoint * handle, const double * dL_dsN, double * dL_ds0,
                                            double * dJ, double * dD, double * dB, double * dK, double * dK_site,
                                            int on_device ) noexcept
try
{
    auto & r = runner( handle );
    DeviceBuffer a, b;
    spins_in( r, dL_dsN, a, on_device );
    b.resize( a.size() );
    r.backward( a.as<double>(), b.as<double>(), dJ, dD, dB, dK, dK_site );
    spins_out( r, b, dL_ds0, on_device );
    return 0;
}
catch( ... )
{
    handle_exception( -1, -1 );
    return 1;
}

extern "C" int Spirit_HIP_Adjoint_Tangent( Spirit_HIP_Adjoint * handle, const double * s0, const double * ds0,
                                           const double * dJ, const double * dD, const double * dB, const double * dK,
                                           const double * dK_site, double * sN, double * dsN, int on_device ) noexcept
try
{
    auto & r = runner( handle );
    DeviceBuffer a, da, b, db;
    spins_in( r, s0, a, on_device );
    spins_in( r, ds0, da, on_device );
    b.resize( a.size() );
    db.resize( a.size() );
    r.tangent( a.as<double>(), da.as<double>(), dJ, dD, dB, dK, dK_site, b.as<double>(), db.as<double>() );
    spins_out( r, b, sN, on_device );
    spins_out( r, db, dsN, on_device );
    return 0;
}
catch( ... )
{
    handle_exception( -1, -1 );
    return 1;
}
