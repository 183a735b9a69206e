// C API: Configurations, Chain, Transitions and Parameters_GNEB.
// Reference: core/src/Spirit/{Configurations,Chain,Transitions,Parameters_GNEB}.cpp,
// core/src/utility/{Configurations,Configuration_Chain}.cpp.
#include "api_common.hpp"

#include <cmath>
#include <functional>

using namespace spirit_hip;

#define API_IMAGE()                                                                                                    \
    std::shared_ptr<Image> image;                                                                                      \
    std::shared_ptr<Chain> chain;                                                                                      \
    from_indices( state, idx_image, idx_chain, image, chain )

using Filter = std::function<bool( const double * pos )>;

// get_filter, core/src/Spirit/Configurations.cpp:12-52 (relative to the geometry centre)
static Filter make_filter( const Image & img, const float position[3], const float r_cut_rectangular[3],
                           float r_cut_cylindrical, float r_cut_spherical, bool inverted )
{
    Vec3 p{ img.geometry->center[0] + position[0], img.geometry->center[1] + position[1],
            img.geometry->center[2] + position[2] };
    float rr[3] = { r_cut_rectangular[0], r_cut_rectangular[1], r_cut_rectangular[2] };
    return [p, rr, r_cut_cylindrical, r_cut_spherical, inverted]( const double * x ) {
        double d0 = x[0] - p[0], d1 = x[1] - p[1], d2 = x[2] - p[2];
        double rc = std::sqrt( std::pow( d0, 2 ) + std::pow( d1, 2 ) );
        double rs = std::sqrt( d0 * d0 + d1 * d1 + d2 * d2 );
        bool in   = ( rr[0] < 0 || std::abs( d0 ) < rr[0] ) && ( rr[1] < 0 || std::abs( d1 ) < rr[1] )
                  && ( rr[2] < 0 || std::abs( d2 ) < rr[2] ) && ( r_cut_cylindrical < 0 || rc < r_cut_cylindrical )
                  && ( r_cut_spherical < 0 || rs < r_cut_spherical );
        return inverted ? !in : in;
    };
}

static void after_spin_change( State * state, int idx_image )
{
    notify_image_changed( state, idx_image, true, false, false );
}

static void normalize3( double * v )
{
    double z = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
    if( z > 0 )
    {
        double n = std::sqrt( z );
        v[0] /= n;
        v[1] /= n;
        v[2] /= n;
    }
}

extern "C" void Configuration_To_Clipboard( State * state, int idx_image, int idx_chain ) noexcept
try
{
    API_IMAGE();
    sync_image_host( state, idx_image );
    state->clipboard_spins = std::make_shared<std::vector<double>>( image->spins );
}
catch( ... )
{
    handle_exception( idx_image, idx_chain );
}

extern "C" void Configuration_From_Clipboard( State * state, const float position[3], const float r_cut_rectangular[3],
                                              float r_cut_cylindrical, float r_cut_spherical, bool inverted,
                                              int idx_image, int idx_chain ) noexcept
try
{
    API_IMAGE();
    if( !state->clipboard_spins || state->clipboard_spins->size() != image->spins.size() )
        throw api_error( 3, "Configuration_From_Clipboard: clipboard empty or of different size" );
    auto f = make_filter( *image, position, r_cut_rectangular, r_cut_cylindrical, r_cut_spherical, inverted );
    std::lock_guard<std::recursive_mutex> lk( image->mtx );
    for( int i = 0; i < image->nos(); ++i )
        if( f( &image->geometry->positions[3 * i] ) )
            for( int c = 0; c < 3; ++c )
                image->spins[3 * i + c] = ( *state->clipboard_spins )[3 * i + c];
    after_spin_change( state, idx_image );
}
catch( ... )
{
    handle_exception( idx_image, idx_chain );
}

// Configurations::Domain, core/src/utility/Configurations.cpp:60-84
extern "C" void Configuration_Domain( State * state, const float direction[3], const float position[3],
                                      const float r_cut_rectangular[3], float r_cut_cylindrical, float r_cut_spherical,
                                      bool inverted, int idx_image, int idx_chain ) noexcept
try
{
    API_IMAGE();
    double v[3] = { direction[0], direction[1], direction[2] };
    if( std::sqrt( v[0] * v[0] + v[1] * v[1] + v[2] * v[2] ) < 1e-8 )
    {
        v[0] = v[1] = 0;
        v[2]        = 1;
    }
    else
        normalize3( v );
    auto f = make_filter( *image, position, r_cut_rectangular, r_cut_cylindrical, r_cut_spherical, inverted );
    std::lock_guard<std::recursive_mutex> lk( image->mtx );
    sync_image_host( state, idx_image );
    for( int i = 0; i < image->nos(); ++i )
        if( f( &image->geometry->positions[3 * i] ) )
            for( int c = 0; c < 3; ++c )
                image->spins[3 * i + c] = v[c];
    after_spin_change( state, idx_image );
}
catch( ... )
{
    handle_exception( idx_image, idx_chain );
}

extern "C" void Configuration_PlusZ( State * state, const float position[3], const float r_cut_rectangular[3],
                                     float r_cut_cylindrical, float r_cut_spherical, bool inverted, int idx_image,
                                     int idx_chain ) noexcept
{
    const float d[3] = { 0, 0, 1 };
    Configuration_Domain( state, d, position, r_cut_rectangular, r_cut_cylindrical, r_cut_spherical, inverted,
                          idx_image, idx_chain );
}

extern "C" void Configuration_MinusZ( State * state, const float position[3], const float r_cut_rectangular[3],
                                      float r_cut_cylindrical, float r_cut_spherical, bool inverted, int idx_image,
                                      int idx_chain ) noexcept
{
    const float d[3] = { 0, 0, -1 };
    Configuration_Domain( state, d, position, r_cut_rectangular, r_cut_cylindrical, r_cut_spherical, inverted,
                          idx_image, idx_chain );
}

// Configurations::Random with get_random_vector_unitsphere (Configurations.cpp:99-126,
// Vectormath.cpp:41-52): the image's mt19937 (seeded by llg_seed) and uniform(-1,1) draws.
extern "C" void Configuration_Random( State * state, const float position[3], const float r_cut_rectangular[3],
                                      float r_cut_cylindrical, float r_cut_spherical, bool inverted, bool external,
                                      int idx_image, int idx_chain ) noexcept
try
{
    API_IMAGE();
    auto f = make_filter( *image, position, r_cut_rectangular, r_cut_cylindrical, r_cut_spherical, inverted );
    std::lock_guard<std::recursive_mutex> lk( image->mtx );
    sync_image_host( state, idx_image );
    std::uniform_real_distribution<double> dist( -1, 1 );
    for( int i = 0; i < image->nos(); ++i )
        if( f( &image->geometry->positions[3 * i] ) )
        {
            double vz              = dist( image->prng );
            double phi             = dist( image->prng ) * Pi;
            double rxy             = std::sqrt( 1 - vz * vz );
            image->spins[3 * i]     = rxy * std::cos( phi );
            image->spins[3 * i + 1] = rxy * std::sin( phi );
            image->spins[3 * i + 2] = vz;
        }
    after_spin_change( state, idx_image );
}
catch( ... )
{
    handle_exception( idx_image, idx_chain );
}

// Configurations::SpinSpiral (Configurations.cpp:162-216), "Real Lattice" / "Reciprocal Lattice"
// direction types with q in units of the lattice; spins rotate around `axis` by theta.
extern "C" void Configuration_SpinSpiral( State * state, const char * direction_type, float q[3], float axis[3],
                                          float theta, const float position[3], const float r_cut_rectangular[3],
                                          float r_cut_cylindrical, float r_cut_spherical, bool inverted,
                                          int idx_image, int idx_chain ) noexcept
try
{
    API_IMAGE();
    auto f = make_filter( *image, position, r_cut_rectangular, r_cut_cylindrical, r_cut_spherical, inverted );
    std::lock_guard<std::recursive_mutex> lk( image->mtx );
    sync_image_host( state, idx_image );
    double th = theta * Pi / 180.0;
    double ax[3] = { axis[0], axis[1], axis[2] };
    normalize3( ax );
    // a frame (e1, e2, ax)
    double e1[3] = { 1, 0, 0 };
    if( std::abs( ax[0] ) > 0.9 )
    {
        e1[0] = 0;
        e1[1] = 1;
    }
    double d = e1[0] * ax[0] + e1[1] * ax[1] + e1[2] * ax[2];
    for( int c = 0; c < 3; ++c )
        e1[c] -= d * ax[c];
    normalize3( e1 );
    double e2[3] = { ax[1] * e1[2] - ax[2] * e1[1], ax[2] * e1[0] - ax[0] * e1[2], ax[0] * e1[1] - ax[1] * e1[0] };
    for( int i = 0; i < image->nos(); ++i )
    {
        const double * x = &image->geometry->positions[3 * i];
        if( !f( x ) )
            continue;
        double phase = 2 * Pi * ( q[0] * x[0] + q[1] * x[1] + q[2] * x[2] ) / image->geometry->lattice_constant;
        for( int c = 0; c < 3; ++c )
            image->spins[3 * i + c] = std::cos( th ) * ax[c]
                                      + std::sin( th ) * ( std::cos( phase ) * e1[c] + std::sin( phase ) * e2[c] );
    }
    after_spin_change( state, idx_image );
}
catch( ... )
{
    handle_exception( idx_image, idx_chain );
}

// Configurations::Add_Noise_Temperature (Configurations.cpp:128-152)
extern "C" void Configuration_Add_Noise_Temperature( State * state, float temperature, const float position[3],
                                                     const float r_cut_rectangular[3], float r_cut_cylindrical,
                                                     float r_cut_spherical, bool inverted, int idx_image,
                                                     int idx_chain ) noexcept
try
{
    API_IMAGE();
    if( temperature == 0.0f )
        return;
    auto f = make_filter( *image, position, r_cut_rectangular, r_cut_cylindrical, r_cut_spherical, inverted );
    std::lock_guard<std::recursive_mutex> lk( image->mtx );
    sync_image_host( state, idx_image );
    double eps = std::sqrt( temperature * k_B );
    std::uniform_real_distribution<double> dist( -1, 1 );
    for( int i = 0; i < image->nos(); ++i )
    {
        double vz  = dist( image->prng );
        double phi = dist( image->prng ) * Pi;
        double rxy = std::sqrt( 1 - vz * vz );
        double xi[3] = { rxy * std::cos( phi ) * eps, rxy * std::sin( phi ) * eps, vz * eps };
        if( f( &image->geometry->positions[3 * i] ) )
            for( int c = 0; c < 3; ++c )
                image->spins[3 * i + c] += xi[c];
        normalize3( &image->spins[3 * i] );
    }
    after_spin_change( state, idx_image );
}
catch( ... )
{
    handle_exception( idx_image, idx_chain );
}

// Configurations::Skyrmion (Configurations.cpp:218-270); the cylindrical cut defaults to r
// (core/src/Spirit/Configurations.cpp:511-538).
extern "C" void Configuration_Skyrmion( State * state, float r, float order, float phase, bool upDown, bool achiral,
                                        bool rl, const float position[3], const float r_cut_rectangular[3],
                                        float r_cut_cylindrical, float r_cut_spherical, bool inverted, int idx_image,
                                        int idx_chain ) noexcept
try
{
    API_IMAGE();
    if( r_cut_cylindrical < 0 )
        r_cut_cylindrical = r;
    auto f = make_filter( *image, position, r_cut_rectangular, r_cut_cylindrical, r_cut_spherical, inverted );
    Vec3 pos{ image->geometry->center[0] + position[0], image->geometry->center[1] + position[1],
              image->geometry->center[2] + position[2] };
    std::lock_guard<std::recursive_mutex> lk( image->mtx );
    sync_image_host( state, idx_image );
    const double r_new = r;
    const int ksi = ( (int)rl ) * 2 - 1, dir = ( (int)upDown ) * 2 - 1;
    auto & P      = image->geometry->positions;
    for( int i = 0; i < image->nos(); ++i )
    {
        double distance = std::sqrt( std::pow( P[3 * i] - pos[0], 2 ) + std::pow( P[3 * i + 1] - pos[1], 2 ) );
        distance        = distance / r_new;
        if( f( &P[3 * i] ) )
        {
            double x     = ( P[3 * i] - pos[0] ) / distance / r_new;
            double phi_i = std::acos( std::max( -1.0, std::min( 1.0, x ) ) );
            if( distance == 0 )
                phi_i = 0;
            if( P[3 * i + 1] - pos[1] < 0.0 )
                phi_i = -phi_i;
            phi_i += phase / 180 * Pi;
            double theta_i          = Pi - Pi * distance;
            image->spins[3 * i]     = ksi * std::sin( theta_i ) * std::cos( order * phi_i );
            image->spins[3 * i + 1] = ksi * std::sin( theta_i ) * std::sin( order * ( phi_i + achiral * Pi ) );
            image->spins[3 * i + 2] = std::cos( theta_i ) * -dir;
        }
    }
    for( int i = 0; i < image->nos(); ++i )
        normalize3( &image->spins[3 * i] );
    after_spin_change( state, idx_image );
}
catch( ... )
{
    handle_exception( idx_image, idx_chain );
}

// Configurations::DW_Skyrmion (Configurations.cpp:272-312): a skyrmion whose profile is a
// pair of tanh domain walls at distance dw_radius from the centre, of width dw_width; the
// cylindrical cut defaults to 3 * max(dw_radius, dw_width) (core/src/Spirit/Configurations.cpp:577-578).
extern "C" void Configuration_DW_Skyrmion( State * state, float dw_radius, float dw_width, float order, float phase,
                                           bool upDown, bool achiral, bool rl, const float position[3],
                                           const float r_cut_rectangular[3], float r_cut_cylindrical,
                                           float r_cut_spherical, bool inverted, int idx_image, int idx_chain ) noexcept
try
{
    API_IMAGE();
    if( r_cut_cylindrical < 0 )
        r_cut_cylindrical = std::max( 3 * dw_radius, 3 * dw_width );
    auto f = make_filter( *image, position, r_cut_rectangular, r_cut_cylindrical, r_cut_spherical, inverted );
    const Vec3 c{ image->geometry->center[0] + position[0], image->geometry->center[1] + position[1],
                  image->geometry->center[2] + position[2] };
    std::lock_guard<std::recursive_mutex> lk( image->mtx );
    sync_image_host( state, idx_image );
    const int ksi = rl ? 1 : -1, dir = upDown ? 1 : -1;
    const double R = dw_radius, W = dw_width;
    auto & P       = image->geometry->positions;
    for( int i = 0; i < image->nos(); ++i )
    {
        const double dx = P[3 * i] - c[0], dy = P[3 * i + 1] - c[1];
        const double r  = std::sqrt( dx * dx + dy * dy );
        if( !f( &P[3 * i] ) )
            continue;
        // polar angle: pi in the core, 0 outside, through the two walls at r = R (and r = -R)
        const double theta = std::asin( std::tanh( -2 * ( r + R ) / W ) ) + std::asin( std::tanh( -2 * ( r - R ) / W ) ) + Pi;
        double phi         = r == 0 ? 0.0 : std::acos( std::max( -1.0, std::min( 1.0, dx / r ) ) );
        if( dy < 0.0 )
            phi = -phi;
        phi += phase / 180 * Pi;
        double * s = &image->spins[3 * i];
        s[0]       = ksi * std::sin( theta ) * std::cos( order * phi );
        s[1]       = ksi * std::sin( theta ) * std::sin( order * phi + achiral * Pi );
        s[2]       = std::cos( theta ) * -dir;
    }
    for( int i = 0; i < image->nos(); ++i )
        normalize3( &image->spins[3 * i] );
    after_spin_change( state, idx_image );
}
catch( ... )
{
    handle_exception( idx_image, idx_chain );
}

// Configurations::Hopfion (Configurations.cpp:156-216): a toroidal hopfion of radius r about
// the axis `normal`, evaluated on the positions turned into the frame whose z axis is the
// normal (Vectormath::dreibein, Vectormath.cpp:410-432) about the centre; the spherical cut
// defaults to pi r (core/src/Spirit/Configurations.cpp:487-488) and is applied to those positions.
extern "C" void Configuration_Hopfion( State * state, float r, int order, const float position[3],
                                       const float r_cut_rectangular[3], float r_cut_cylindrical,
                                       float r_cut_spherical, bool inverted, const float normal[3], int idx_image,
                                       int idx_chain ) noexcept
try
{
    API_IMAGE();
    if( r_cut_spherical < 0 )
        r_cut_spherical = r * (float)Pi;
    auto f = make_filter( *image, position, r_cut_rectangular, r_cut_cylindrical, r_cut_spherical, inverted );
    const Vec3 c{ image->geometry->center[0] + position[0], image->geometry->center[1] + position[1],
                  image->geometry->center[2] + position[2] };
    // frame: rows x_hat, y_hat, n (normalised); x_hat = n x e_x unless n is close to e_x
    static const float default_normal[3] = { 0, 0, 1 };
    if( !normal )
        normal = default_normal;
    double n[3] = { normal[0], normal[1], normal[2] };
    const double ex[3] = { 1, 0, 0 }, ey[3] = { 0, 1, 0 };
    auto cross = []( const double * a, const double * b, double * o ) {
        o[0] = a[1] * b[2] - a[2] * b[1], o[1] = a[2] * b[0] - a[0] * b[2], o[2] = a[0] * b[1] - a[1] * b[0];
    };
    double xh[3], yh[3];
    cross( n, std::abs( n[0] - 1 ) > 1e-1 ? ex : ey, xh );
    cross( n, xh, yh );
    normalize3( xh );
    normalize3( yh );
    normalize3( n );
    const double Rm[3][3] = { { xh[0], xh[1], xh[2] }, { yh[0], yh[1], yh[2] }, { n[0], n[1], n[2] } };
    std::lock_guard<std::recursive_mutex> lk( image->mtx );
    sync_image_host( state, idx_image );
    auto & P = image->geometry->positions;
    if( r != 0 )
        for( int i = 0; i < image->nos(); ++i )
        {
            // turned position about the centre
            const double d0 = P[3 * i] - c[0], d1 = P[3 * i + 1] - c[1], d2 = P[3 * i + 2] - c[2];
            double q[3];
            for( int k = 0; k < 3; ++k )
                q[k] = Rm[k][0] * d0 + Rm[k][1] * d1 + Rm[k][2] * d2 + c[k];
            if( !f( q ) )
                continue;
            const double dq[3] = { q[0] - c[0], q[1] - c[1], q[2] - c[2] };
            const double d     = std::sqrt( dq[0] * dq[0] + dq[1] * dq[1] + dq[2] * dq[2] );
            // polar angle against the torus axis, and the profile of the tube
            const double T   = std::acos( d == 0 ? 0.0 : dq[2] / d );
            double t         = d / r;
            t                = 1.0 + 4.22 / ( t * t );
            const double tmp = Pi * ( 1.0 - 1.0 / std::sqrt( t ) );
            t                = std::sin( tmp ) * std::sin( T );
            t                = std::acos( 1.0 - 2.0 * t * t );
            const double F   = std::atan2( dq[1], dq[0] );
            const double fr  = F + std::atan( 1.0 / ( std::tan( tmp ) * std::cos( T ) ) ) + ( T > Pi / 2.0 ? 0.0 : Pi );
            double * s       = &image->spins[3 * i];
            s[0]             = std::sin( t ) * std::cos( order * fr );
            s[1]             = std::sin( t ) * std::sin( order * fr );
            s[2]             = std::cos( t );
        }
    after_spin_change( state, idx_image );
}
catch( ... )
{
    handle_exception( idx_image, idx_chain );
}

// ===========================================================================
// Chain.h  (core/src/Spirit/Chain.cpp)
// ===========================================================================
static void chain_setup_data( State * state )
{
    auto & c = *state->chain;
    c.Rx.assign( c.noi(), 0.0 );
    int n = ( c.noi() - 1 ) * c.gneb->n_E_interpolations + c.noi();
    c.Rx_interpolated.assign( std::max( 0, n ), 0.0 );
    c.E_interpolated.assign( std::max( 0, n ), 0.0 );
    c.image_type.resize( c.noi(), GNEB_NORMAL );
}

extern "C" int Chain_Get_NOI( State * state, int idx_chain ) noexcept
try
{
    return state->chain->noi();
}
catch( ... )
{
    handle_exception( -1, idx_chain );
    return 0;
}

extern "C" bool Chain_Jump_To_Image( State * state, int idx_image, int idx_chain ) noexcept
try
{
    API_IMAGE();
    state->idx_active_image = idx_image;
    State_Update( state );
    return true;
}
catch( ... )
{
    handle_exception( idx_image, idx_chain );
    return false;
}

extern "C" bool Chain_next_Image( State * state, int idx_chain ) noexcept
try
{
    if( state->idx_active_image < state->chain->noi() - 1 )
    {
        ++state->idx_active_image;
        State_Update( state );
        return true;
    }
    return false;
}
catch( ... )
{
    handle_exception( -1, idx_chain );
    return false;
}

extern "C" bool Chain_prev_Image( State * state, int idx_chain ) noexcept
try
{
    if( state->idx_active_image > 0 )
    {
        --state->idx_active_image;
        State_Update( state );
        return true;
    }
    return false;
}
catch( ... )
{
    handle_exception( -1, idx_chain );
    return false;
}

extern "C" void Chain_Image_to_Clipboard( State * state, int idx_image, int idx_chain ) noexcept
try
{
    API_IMAGE();
    sync_image_host( state, idx_image );
    state->clipboard_image = std::make_shared<Image>( *image );
}
catch( ... )
{
    handle_exception( idx_image, idx_chain );
}

extern "C" void Chain_Replace_Image( State * state, int idx_image, int idx_chain ) noexcept
try
{
    API_IMAGE();
    if( !state->clipboard_image )
        throw api_error( 3, "Tried to replace image, but clipboard was empty." );
    if( image->iteration_allowed || chain->iteration_allowed )
        Simulation_Stop( state, idx_image, idx_chain );
    chain->images[idx_image] = std::make_shared<Image>( *state->clipboard_image );
    state->method_image.erase( idx_image );
    State_Update( state );
}
catch( ... )
{
    handle_exception( idx_image, idx_chain );
}

static void insert_image( State * state, int pos )
{
    auto & c = *state->chain;
    if( c.iteration_allowed )
        Simulation_Stop( state, -1, -1 );
    auto copy = std::make_shared<Image>( *state->clipboard_image );
    c.images.insert( c.images.begin() + pos, copy );
    c.image_type.insert( c.image_type.begin() + pos, GNEB_NORMAL );
    // image-indexed methods shift
    std::map<int, std::shared_ptr<MethodLLG>> m;
    for( auto & kv : state->method_image )
        m[kv.first >= pos ? kv.first + 1 : kv.first] = kv.second;
    state->method_image = m;
    if( state->idx_active_image >= pos )
        ++state->idx_active_image;
    State_Update( state );
    chain_setup_data( state );
}

extern "C" void Chain_Insert_Image_Before( State * state, int idx_image, int idx_chain ) noexcept
try
{
    API_IMAGE();
    if( !state->clipboard_image )
        throw api_error( 3, "Tried to insert image before, but clipboard was empty." );
    insert_image( state, idx_image );
}
catch( ... )
{
    handle_exception( idx_image, idx_chain );
}

extern "C" void Chain_Insert_Image_After( State * state, int idx_image, int idx_chain ) noexcept
try
{
    API_IMAGE();
    if( !state->clipboard_image )
        throw api_error( 3, "Tried to insert image after, but clipboard was empty." );
    insert_image( state, idx_image + 1 );
}
catch( ... )
{
    handle_exception( idx_image, idx_chain );
}

extern "C" void Chain_Push_Back( State * state, int idx_chain ) noexcept
try
{
    if( !state->clipboard_image )
        throw api_error( 3, "Tried to push back image, but clipboard was empty." );
    insert_image( state, state->chain->noi() );
}
catch( ... )
{
    handle_exception( -1, idx_chain );
}

extern "C" bool Chain_Delete_Image( State * state, int idx_image, int idx_chain ) noexcept
try
{
    API_IMAGE();
    auto & c = *state->chain;
    if( c.noi() <= 1 )
        throw api_error( 3, "Tried to delete last image." );
    if( image->iteration_allowed || c.iteration_allowed )
        Simulation_Stop( state, idx_image, idx_chain );
    c.images.erase( c.images.begin() + idx_image );
    c.image_type.erase( c.image_type.begin() + idx_image );
    std::map<int, std::shared_ptr<MethodLLG>> m;
    for( auto & kv : state->method_image )
        if( kv.first != idx_image )
            m[kv.first > idx_image ? kv.first - 1 : kv.first] = kv.second;
    state->method_image = m;
    if( state->idx_active_image > idx_image || state->idx_active_image >= c.noi() )
        --state->idx_active_image;
    if( state->idx_active_image < 0 )
        state->idx_active_image = 0;
    State_Update( state );
    chain_setup_data( state );
    return true;
}
catch( ... )
{
    handle_exception( idx_image, idx_chain );
    return false;
}

extern "C" bool Chain_Pop_Back( State * state, int idx_chain ) noexcept
{
    return Chain_Delete_Image( state, state->chain->noi() - 1, idx_chain );
}

extern "C" void Chain_Set_Length( State * state, int n_images, int idx_chain ) noexcept
try
{
    if( n_images < 1 )
        throw api_error( 3, "Tried to set chain length to less than 1." );
    if( !state->clipboard_image )
        state->clipboard_image = std::make_shared<Image>( *state->active_image );
    while( state->chain->noi() < n_images )
        insert_image( state, state->chain->noi() );
    while( state->chain->noi() > n_images )
        Chain_Pop_Back( state, idx_chain );
}
catch( ... )
{
    handle_exception( -1, idx_chain );
}

extern "C" void Chain_Get_Rx( State * state, float * Rx, int idx_chain ) noexcept
try
{
    for( size_t i = 0; i < state->chain->Rx.size(); ++i )
        Rx[i] = (float)state->chain->Rx[i];
}
catch( ... )
{
    handle_exception( -1, idx_chain );
}

extern "C" void Chain_Get_Rx_Interpolated( State * state, float * Rx_interpolated, int idx_chain ) noexcept
try
{
    for( size_t i = 0; i < state->chain->Rx_interpolated.size(); ++i )
        Rx_interpolated[i] = (float)state->chain->Rx_interpolated[i];
}
catch( ... )
{
    handle_exception( -1, idx_chain );
}

extern "C" void Chain_Get_Energy( State * state, float * energy, int idx_chain ) noexcept
try
{
    for( int i = 0; i < state->chain->noi(); ++i )
        energy[i] = (float)state->chain->images[i]->E;
}
catch( ... )
{
    handle_exception( -1, idx_chain );
}

extern "C" void Chain_Get_Energy_Interpolated( State * state, float * E_interpolated, int idx_chain ) noexcept
try
{
    for( size_t i = 0; i < state->chain->E_interpolated.size(); ++i )
        E_interpolated[i] = (float)state->chain->E_interpolated[i];
}
catch( ... )
{
    handle_exception( -1, idx_chain );
}

extern "C" void Chain_Update_Data( State * state, int idx_chain ) noexcept;
extern "C" void Chain_Setup_Data( State * state, int idx_chain ) noexcept
try
{
    chain_setup_data( state );
}
catch( ... )
{
    handle_exception( -1, idx_chain );
}

// ===========================================================================
// Transitions.h  (core/src/utility/Configuration_Chain.cpp:20-75)
// ===========================================================================
static double angle3( const double * a, const double * b )
{
    double r = a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
    r        = std::fmax( -1.0, std::fmin( 1.0, r ) );
    return std::acos( r );
}

static void rotate3( const double * v, const double * axis, double angle, double * out )
{
    double c = std::cos( angle ), s = std::sin( angle );
    double cr[3] = { axis[1] * v[2] - axis[2] * v[1], axis[2] * v[0] - axis[0] * v[2], axis[0] * v[1] - axis[1] * v[0] };
    double d     = axis[0] * v[0] + axis[1] * v[1] + axis[2] * v[2];
    double o[3];
    for( int q = 0; q < 3; ++q )
        o[q] = v[q] * c + cr[q] * s + axis[q] * d * ( 1 - c );
    out[0] = o[0], out[1] = o[1], out[2] = o[2];
}

extern "C" void Transition_Homogeneous( State * state, int idx_1, int idx_2, int idx_chain ) noexcept
try
{
    auto & c = *state->chain;
    if( idx_1 < 0 || idx_2 >= c.noi() || idx_1 >= idx_2 )
        throw api_error( 3, "Transition_Homogeneous: invalid image indices" );
    if( c.iteration_allowed )
        throw api_error( 3, "Transition_Homogeneous: a GNEB simulation is running" );
    sync_all_hosts( state );
    auto & s1 = c.images[idx_1]->spins;
    auto & s2 = c.images[idx_2]->spins;
    const int nos = c.images[0]->nos();
    const double ex[3] = { 1, 0, 0 }, ey[3] = { 0, 1, 0 };
    bool antiparallel = false;
    for( int i = 0; i < nos; ++i )
    {
        const double * a = &s1[3 * i];
        const double * b = &s2[3 * i];
        double rot_angle = angle3( a, b );
        double axis[3]   = { a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0] };
        normalize3( axis );
        if( std::abs( rot_angle - Pi ) < 1e-4 )
        {
            antiparallel = true;
            double d     = a[0] * ex[0] + a[1] * ex[1] + a[2] * ex[2];
            const double * ax = ( std::abs( d ) - 1 > 1e-4 ) ? ex : ey;
            axis[0] = ax[0], axis[1] = ax[1], axis[2] = ax[2];
        }
        for( int img = idx_1 + 1; img < idx_2; ++img )
        {
            double * o = &c.images[img]->spins[3 * i];
            if( rot_angle > 1e-8 )
            {
                double ang = rot_angle * double( img - idx_1 ) / double( idx_2 - idx_1 );
                rotate3( a, axis, ang, o );
            }
            else
                o[0] = a[0], o[1] = a[1], o[2] = a[2];
        }
    }
    if( antiparallel )
        log_message( 3, "For the interpolation of antiparallel spins an arbitrary rotation axis has been chosen." );
    for( int img = idx_1 + 1; img < idx_2; ++img )
        notify_image_changed( state, img, true, false, false );
}
catch( ... )
{
    handle_exception( -1, idx_chain );
}

extern "C" void Transition_Homogeneous_Insert_Interpolated( State * state, int n_interpolate, int idx_chain ) noexcept
try
{
    if( n_interpolate <= 0 )
        return;
    int noi = state->chain->noi();
    for( int img = noi - 2; img >= 0; --img )
    {
        Chain_Image_to_Clipboard( state, img, idx_chain );
        for( int k = 0; k < n_interpolate; ++k )
            insert_image( state, img + 1 );
        Transition_Homogeneous( state, img, img + n_interpolate + 1, idx_chain );
    }
}
catch( ... )
{
    handle_exception( -1, idx_chain );
}

extern "C" void Transition_Add_Noise_Temperature( State * state, float temperature, int idx_1, int idx_2,
                                                  int idx_chain ) noexcept
try
{
    const float pos[3] = { 0, 0, 0 }, rect[3] = { -1, -1, -1 };
    for( int img = idx_1 + 1; img <= idx_2 - 1; ++img )
        Configuration_Add_Noise_Temperature( state, temperature, pos, rect, -1, -1, false, img, idx_chain );
}
catch( ... )
{
    handle_exception( -1, idx_chain );
}

// ===========================================================================
// Parameters_GNEB.h  (core/src/Spirit/Parameters_GNEB.cpp)
// ===========================================================================
// every setter reaches a running single-shot GNEB method when it returns
struct GnebNotify
{
    State * s;
    ~GnebNotify()
    {
        gneb_notify_changed( s );
    }
};
#define GNEB_PARAMS                                                                                                    \
    auto & p = *state->chain->gneb;                                                                                    \
    GnebNotify gneb_notify_on_return{ state }

extern "C" void Parameters_GNEB_Set_Output_Tag( State * state, const char * tag, int idx_chain ) noexcept
{
    GNEB_PARAMS;
    p.output_file_tag = tag;
}
extern "C" void Parameters_GNEB_Set_Output_Folder( State * state, const char * folder, int idx_chain ) noexcept
{
    GNEB_PARAMS;
    p.output_folder = folder;
}
extern "C" void Parameters_GNEB_Set_Output_General( State * state, bool any, bool initial, bool final,
                                                    int idx_chain ) noexcept
{
    GNEB_PARAMS;
    p.output_any     = any;
    p.output_initial = initial;
    p.output_final   = final;
}
extern "C" void Parameters_GNEB_Set_N_Iterations( State * state, int n_iterations, int n_iterations_log,
                                                  int idx_chain ) noexcept
{
    GNEB_PARAMS;
    p.n_iterations     = n_iterations;
    p.n_iterations_log = n_iterations_log;
}
extern "C" void Parameters_GNEB_Set_Convergence( State * state, float convergence, int idx_image, int idx_chain ) noexcept
{
    GNEB_PARAMS;
    p.force_convergence = convergence;
}
extern "C" void Parameters_GNEB_Set_Spring_Constant( State * state, float spring_constant, int idx_image,
                                                     int idx_chain ) noexcept
{
    GNEB_PARAMS;
    p.spring_constant = spring_constant;
}
extern "C" void Parameters_GNEB_Set_Spring_Force_Ratio( State * state, float ratio, int idx_chain ) noexcept
{
    GNEB_PARAMS;
    p.spring_force_ratio = ratio;
}
extern "C" void Parameters_GNEB_Set_Path_Shortening_Constant( State * state, float path_shortening_constant,
                                                              int idx_chain ) noexcept
{
    GNEB_PARAMS;
    p.path_shortening_constant = path_shortening_constant;
}
extern "C" void Parameters_GNEB_Set_Moving_Endpoints( State * state, bool moving_endpoints, int idx_chain ) noexcept
{
    GNEB_PARAMS;
    p.moving_endpoints = moving_endpoints;
}
extern "C" void Parameters_GNEB_Set_Translating_Endpoints( State * state, bool translating_endpoints,
                                                           int idx_chain ) noexcept
{
    GNEB_PARAMS;
    p.translating_endpoints = translating_endpoints;
}
extern "C" void Parameters_GNEB_Set_Equilibrium_Delta_Rx( State * state, float delta_Rx_left, float delta_Rx_right,
                                                          int idx_chain ) noexcept
{
    GNEB_PARAMS;
    p.equilibrium_delta_Rx_left  = delta_Rx_left;
    p.equilibrium_delta_Rx_right = delta_Rx_right;
}
extern "C" void Parameters_GNEB_Set_Climbing_Falling( State * state, int image_type, int idx_image,
                                                      int idx_chain ) noexcept
try
{
    API_IMAGE();
    chain->image_type.resize( chain->noi(), GNEB_NORMAL );
    chain->image_type[idx_image] = image_type;
    gneb_notify_changed( state );
}
catch( ... )
{
    handle_exception( idx_image, idx_chain );
}

// Parameters_GNEB.cpp:338-363: local energy maxima climb, minima fall, stationary images stay
extern "C" void Parameters_GNEB_Set_Image_Type_Automatically( State * state, int idx_chain ) noexcept
try
{
    auto & c = *state->chain;
    c.image_type.resize( c.noi(), GNEB_NORMAL );
    for( int img = 1; img < c.noi() - 1; ++img )
    {
        double E0 = c.images[img - 1]->E, E1 = c.images[img]->E, E2 = c.images[img + 1]->E;
        if( E0 < E1 && E1 > E2 )
            c.image_type[img] = GNEB_CLIMBING;
        else if( E0 > E1 && E1 < E2 )
            c.image_type[img] = GNEB_FALLING;
        else if( c.image_type[img] != GNEB_STATIONARY )
            c.image_type[img] = GNEB_NORMAL;
    }
    gneb_notify_changed( state );
}
catch( ... )
{
    handle_exception( -1, idx_chain );
}

extern "C" void Parameters_GNEB_Set_N_Energy_Interpolations( State * state, int n, int idx_chain ) noexcept
{
    GNEB_PARAMS;
    p.n_E_interpolations = n;
    chain_setup_data( state );
}
extern "C" const char * Parameters_GNEB_Get_Output_Tag( State * state, int idx_chain ) noexcept
{
    return state->chain->gneb->output_file_tag.c_str();
}
extern "C" const char * Parameters_GNEB_Get_Output_Folder( State * state, int idx_chain ) noexcept
{
    return state->chain->gneb->output_folder.c_str();
}
extern "C" void Parameters_GNEB_Get_Output_General( State * state, bool * any, bool * initial, bool * final,
                                                    int idx_chain ) noexcept
{
    GNEB_PARAMS;
    *any     = p.output_any;
    *initial = p.output_initial;
    *final   = p.output_final;
}
extern "C" void Parameters_GNEB_Get_N_Iterations( State * state, int * iterations, int * iterations_log,
                                                  int idx_chain ) noexcept
{
    GNEB_PARAMS;
    *iterations     = (int)p.n_iterations;
    *iterations_log = (int)p.n_iterations_log;
}
extern "C" float Parameters_GNEB_Get_Convergence( State * state, int idx_image, int idx_chain ) noexcept
{
    return (float)state->chain->gneb->force_convergence;
}
extern "C" float Parameters_GNEB_Get_Spring_Constant( State * state, int idx_image, int idx_chain ) noexcept
{
    return (float)state->chain->gneb->spring_constant;
}
extern "C" float Parameters_GNEB_Get_Spring_Force_Ratio( State * state, int idx_chain ) noexcept
{
    return (float)state->chain->gneb->spring_force_ratio;
}
extern "C" float Parameters_GNEB_Get_Path_Shortening_Constant( State * state, int idx_chain ) noexcept
{
    return (float)state->chain->gneb->path_shortening_constant;
}
extern "C" bool Parameters_GNEB_Get_Moving_Endpoints( State * state, int idx_chain ) noexcept
{
    return state->chain->gneb->moving_endpoints;
}
extern "C" bool Parameters_GNEB_Get_Translating_Endpoints( State * state, int idx_chain ) noexcept
{
    return state->chain->gneb->translating_endpoints;
}
extern "C" void Parameters_GNEB_Get_Equilibrium_Delta_Rx( State * state, float * delta_Rx_left, float * delta_Rx_right,
                                                          int idx_chain ) noexcept
{
    GNEB_PARAMS;
    *delta_Rx_left  = (float)p.equilibrium_delta_Rx_left;
    *delta_Rx_right = (float)p.equilibrium_delta_Rx_right;
}
extern "C" int Parameters_GNEB_Get_Climbing_Falling( State * state, int idx_image, int idx_chain ) noexcept
try
{
    API_IMAGE();
    return chain->image_type[idx_image];
}
catch( ... )
{
    handle_exception( idx_image, idx_chain );
    return 0;
}
extern "C" int Parameters_GNEB_Get_N_Energy_Interpolations( State * state, int idx_chain ) noexcept
{
    return state->chain->gneb->n_E_interpolations;
}
