// C API: Hamiltonian, System, Geometry, Parameters, Quantities, Constants, Version, Log
// and the HIP-only extensions.  Reference: core/src/Spirit/{Hamiltonian,System,Geometry,
// Parameters_LLG,Parameters_GNEB,Quantities,Constants,Version,Log}.cpp.
#include "api_common.hpp"
#include "hessian.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>

using namespace spirit_hip;

#define API_IMAGE()                                                                                                    \
    std::shared_ptr<Image> image;                                                                                      \
    std::shared_ptr<Chain> chain;                                                                                      \
    from_indices( state, idx_image, idx_chain, image, chain )

static Vec3 as_vec3( const float * v )
{
    return Vec3{ v[0], v[1], v[2] };
}

// Parameters_LLG_Set_STT (Parameters_LLG.cpp:324-338): a normal shorter than 0.9 becomes (0, 0, 1)
static Vec3 stt_normal( const float * v )
{
    const double len2 = (double)v[0] * v[0] + (double)v[1] * v[1] + (double)v[2] * v[2];
    if( len2 < 0.81 )
    {
        log_message( 3, "s_c_vec = {0,0,0} replaced by {0,0,1}" );
        return Vec3{ 0, 0, 1 };
    }
    Vec3 n{ v[0], v[1], v[2] };
    const double s = std::sqrt( len2 );
    for( auto & x : n )
        x /= s;
    return n;
}

static Vec3 normalized3( const float * v )
{
    Vec3 n{ v[0], v[1], v[2] };
    double z = n[0] * n[0] + n[1] * n[1] + n[2] * n[2];
    if( z > 0 )
    {
        double s = std::sqrt( z );
        for( auto & x : n )
            x /= s;
    }
    return n;
}

// ===========================================================================
// Hamiltonian.h  (core/src/Spirit/Hamiltonian.cpp)
// ===========================================================================
extern "C" void Hamiltonian_Set_Boundary_Conditions( State * state, const bool * periodical, int idx_image,
                                                     int idx_chain ) noexcept
try
{
    API_IMAGE();
    std::lock_guard<std::recursive_mutex> lk( image->mtx );
    for( int c = 0; c < 3; ++c )
        image->hamiltonian->bc[c] = periodical[c] ? 1 : 0;
    ++image->hamiltonian->version;
    notify_image_changed( state, idx_image, false, true, false );
}
catch( ... )
{
    handle_exception( idx_image, idx_chain );
}

extern "C" void Hamiltonian_Set_Field( State * state, float magnitude, const float * normal, int idx_image,
                                       int idx_chain ) noexcept
try
{
    API_IMAGE();
    std::lock_guard<std::recursive_mutex> lk( image->mtx );
    image->hamiltonian->field_normal    = normalized3( normal );
    image->hamiltonian->field_magnitude = magnitude * mu_B;
    ++image->hamiltonian->version;
    notify_image_changed( state, idx_image, false, true, false );
}
catch( ... )
{
    handle_exception( idx_image, idx_chain );
}

extern "C" void Hamiltonian_Set_Anisotropy( State * state, float magnitude, const float * normal, int idx_image,
                                            int idx_chain ) noexcept
try
{
    API_IMAGE();
    std::lock_guard<std::recursive_mutex> lk( image->mtx );
    auto & h = *image->hamiltonian;
    int nb   = image->geometry->nb;
    Vec3 n   = normalized3( normal );
    h.aniso_idx.assign( nb, 0 );
    h.aniso_K.assign( nb, magnitude );
    h.aniso_n.assign( nb, n );
    for( int i = 0; i < nb; ++i )
        h.aniso_idx[i] = i;
    ++h.version;
    notify_image_changed( state, idx_image, false, true, false );
}
catch( ... )
{
    handle_exception( idx_image, idx_chain );
}

extern "C" void Hamiltonian_Set_Cubic_Anisotropy( State * state, float magnitude, int idx_image, int idx_chain ) noexcept
try
{
    API_IMAGE();
    std::lock_guard<std::recursive_mutex> lk( image->mtx );
    auto & h = *image->hamiltonian;
    int nb   = image->geometry->nb;
    h.cubic_idx.assign( nb, 0 );
    h.cubic_K.assign( nb, magnitude );
    for( int i = 0; i < nb; ++i )
        h.cubic_idx[i] = i;
    ++h.version;
    notify_image_changed( state, idx_image, false, true, false );
}
catch( ... )
{
    handle_exception( idx_image, idx_chain );
}

extern "C" void Hamiltonian_Set_Exchange( State * state, int n_shells, const float * jij, int idx_image,
                                          int idx_chain ) noexcept
try
{
    API_IMAGE();
    std::lock_guard<std::recursive_mutex> lk( image->mtx );
    auto & h = *image->hamiltonian;
    h.exchange_shell_J.assign( jij, jij + std::max( 0, n_shells ) );
    h.exch_pairs_in.clear();
    h.exch_J_in.clear();
    h.update_interactions( *image->geometry );
    notify_image_changed( state, idx_image, false, true, false );
}
catch( ... )
{
    handle_exception( idx_image, idx_chain );
}

extern "C" void Hamiltonian_Set_DMI( State * state, int n_shells, const float * dij, int chirality, int idx_image,
                                     int idx_chain ) noexcept
try
{
    API_IMAGE();
    if( chirality != 1 && chirality != 2 && chirality != -1 && chirality != -2 )
        throw api_error( 2, "Hamiltonian_Set_DMI: invalid chirality " + std::to_string( chirality ) );
    std::lock_guard<std::recursive_mutex> lk( image->mtx );
    auto & h = *image->hamiltonian;
    h.dmi_shell_D.assign( dij, dij + std::max( 0, n_shells ) );
    h.dmi_chirality = chirality;
    h.dmi_pairs_in.clear();
    h.dmi_D_in.clear();
    h.dmi_n_in.clear();
    h.update_interactions( *image->geometry );
    notify_image_changed( state, idx_image, false, true, false );
}
catch( ... )
{
    handle_exception( idx_image, idx_chain );
}

extern "C" void Hamiltonian_Set_DDI( State * state, int ddi_method, int n_periodic_images[3], float cutoff_radius,
                                     bool pb_zero_padding, int idx_image, int idx_chain ) noexcept
try
{
    API_IMAGE();
    std::lock_guard<std::recursive_mutex> lk( image->mtx );
    auto & h              = *image->hamiltonian;
    h.ddi_method          = ddi_method;
    h.ddi_images          = { n_periodic_images[0], n_periodic_images[1], n_periodic_images[2] };
    h.ddi_radius          = cutoff_radius;
    h.ddi_pb_zero_padding = pb_zero_padding;
    if( ddi_method == DDI_FMM )
        log_message( 2, "Hamiltonian_Set_DDI: the FMM method has no dipolar field (as Gradient_DDI, "
                        "Hamiltonian_Heisenberg.cpp:866-878, evaluates none for it)" );
    h.update_interactions( *image->geometry );
    notify_image_changed( state, idx_image, false, true, false );
}
catch( ... )
{
    handle_exception( idx_image, idx_chain );
}

extern "C" const char * Hamiltonian_Get_Name( State * state, int idx_image, int idx_chain ) noexcept
{
    return "Heisenberg";
}

extern "C" void Hamiltonian_Get_Boundary_Conditions( State * state, bool * periodical, int idx_image,
                                                     int idx_chain ) noexcept
try
{
    API_IMAGE();
    for( int c = 0; c < 3; ++c )
        periodical[c] = image->hamiltonian->bc[c] != 0;
}
catch( ... )
{
    handle_exception( idx_image, idx_chain );
}

extern "C" void Hamiltonian_Get_Field( State * state, float * magnitude, float * normal, int idx_image,
                                       int idx_chain ) noexcept
try
{
    API_IMAGE();
    auto & h = *image->hamiltonian;
    if( h.field_magnitude > 0 )
    {
        *magnitude = (float)( h.field_magnitude / mu_B );
        for( int c = 0; c < 3; ++c )
            normal[c] = (float)h.field_normal[c];
    }
    else
    {
        *magnitude = 0;
        normal[0] = normal[1] = 0;
        normal[2]             = 1;
    }
}
catch( ... )
{
    handle_exception( idx_image, idx_chain );
}

extern "C" void Hamiltonian_Get_Anisotropy( State * state, float * magnitude, float * normal, int idx_image,
                                            int idx_chain ) noexcept
try
{
    API_IMAGE();
    auto & h = *image->hamiltonian;
    if( !h.aniso_idx.empty() )
    {
        *magnitude = (float)h.aniso_K[0];
        for( int c = 0; c < 3; ++c )
            normal[c] = (float)h.aniso_n[0][c];
    }
    else
    {
        *magnitude = 0;
        normal[0] = normal[1] = 0;
        normal[2]             = 1;
    }
}
catch( ... )
{
    handle_exception( idx_image, idx_chain );
}

extern "C" void Hamiltonian_Get_Cubic_Anisotropy( State * state, float * magnitude, int idx_image, int idx_chain ) noexcept
try
{
    API_IMAGE();
    *magnitude = image->hamiltonian->cubic_K.empty() ? 0.f : (float)image->hamiltonian->cubic_K[0];
}
catch( ... )
{
    handle_exception( idx_image, idx_chain );
}

extern "C" void Hamiltonian_Get_Exchange_Shells( State * state, int * n_shells, float * jij, int idx_image,
                                                 int idx_chain ) noexcept
try
{
    API_IMAGE();
    auto & h  = *image->hamiltonian;
    *n_shells = (int)h.exchange_shell_J.size();
    if( jij )
        for( size_t i = 0; i < h.exchange_shell_J.size(); ++i )
            jij[i] = (float)h.exchange_shell_J[i];
}
catch( ... )
{
    handle_exception( idx_image, idx_chain );
}

extern "C" int Hamiltonian_Get_Exchange_N_Pairs( State * state, int idx_image, int idx_chain ) noexcept
try
{
    API_IMAGE();
    return (int)image->hamiltonian->exch_pairs.size();
}
catch( ... )
{
    handle_exception( idx_image, idx_chain );
    return 0;
}

extern "C" void Hamiltonian_Get_Exchange_Pairs( State * state, int idx[][2], int translations[][3], float * Jij,
                                                int idx_image, int idx_chain ) noexcept
try
{
    API_IMAGE();
    auto & h = *image->hamiltonian;
    for( size_t k = 0; k < h.exch_pairs.size(); ++k )
    {
        idx[k][0] = h.exch_pairs[k].i;
        idx[k][1] = h.exch_pairs[k].j;
        for( int c = 0; c < 3; ++c )
            translations[k][c] = h.exch_pairs[k].t[c];
        Jij[k] = (float)h.exch_J[k];
    }
}
catch( ... )
{
    handle_exception( idx_image, idx_chain );
}

extern "C" void Hamiltonian_Get_DMI_Shells( State * state, int * n_shells, float * dij, int * chirality,
                                            int idx_image, int idx_chain ) noexcept
try
{
    API_IMAGE();
    auto & h  = *image->hamiltonian;
    *n_shells = (int)h.dmi_shell_D.size();
    if( dij )
        for( size_t i = 0; i < h.dmi_shell_D.size(); ++i )
            dij[i] = (float)h.dmi_shell_D[i];
    if( chirality )
        *chirality = h.dmi_chirality;
}
catch( ... )
{
    handle_exception( idx_image, idx_chain );
}

extern "C" int Hamiltonian_Get_DMI_N_Pairs( State * state, int idx_image, int idx_chain ) noexcept
try
{
    API_IMAGE();
    return (int)image->hamiltonian->dmi_pairs.size();
}
catch( ... )
{
    handle_exception( idx_image, idx_chain );
    return 0;
}

extern "C" void Hamiltonian_Get_DDI( State * state, int * ddi_method, int n_periodic_images[3], float * cutoff_radius,
                                     bool * pb_zero_padding, int idx_image, int idx_chain ) noexcept
try
{
    API_IMAGE();
    auto & h    = *image->hamiltonian;
    *ddi_method = h.ddi_method;
    for( int c = 0; c < 3; ++c )
        n_periodic_images[c] = h.ddi_images[c];
    *cutoff_radius   = (float)h.ddi_radius;
    *pb_zero_padding = h.ddi_pb_zero_padding;
}
catch( ... )
{
    handle_exception( idx_image, idx_chain );
}

// The embedding Hessian as triplets "row<TAB>col<TAB>value" per line, column-major, or as the
// dense 3N x 3N matrix with one row per line, "value " per column and "0 " where the sparse
// matrix holds no entry (core/src/Spirit/Hamiltonian.cpp:672-727 with Eigen's sparse-matrix
// stream output, core/thirdparty/Eigen/src/SparseCore/SparseMatrixBase.h:218-262). Values
// are printed as a default C++ stream prints a double (six significant digits, %g).
extern "C" void Hamiltonian_Write_Hessian( State * state, const char * filename, bool triplet_format, int idx_image,
                                           int idx_chain ) noexcept
try
{
    API_IMAGE();
    if( filename == nullptr )
        throw std::runtime_error( "Hamiltonian_Write_Hessian: no file name given" );
    const auto entries = embedding_hessian( *image->geometry, *image->hamiltonian );
    const std::string fname( filename );
    std::printf( "Saving %s to file: %s\n", triplet_format ? "triplets" : "matrix", fname.c_str() );
    std::fflush( stdout );
    std::FILE * f = std::fopen( fname.c_str(), "w" );
    if( f == nullptr )
    {
        std::fprintf( stderr, "Could not save matrix!" );
        return;
    }
    char num[64];
    if( triplet_format )
    {
        for( const auto & e : entries )
        {
            std::snprintf( num, sizeof( num ), "%g", e.value );
            std::fprintf( f, "%ld\t%ld\t%s\n", e.row, e.col, num );
        }
    }
    else
    {
        // row-major walk over the column-major entries
        const long n = 3L * image->nos();
        std::vector<HessianEntry> rows( entries );
        std::stable_sort( rows.begin(), rows.end(), []( const HessianEntry & x, const HessianEntry & y ) {
            return x.row != y.row ? x.row < y.row : x.col < y.col;
        } );
        size_t k = 0;
        for( long r = 0; r < n; ++r )
        {
            long c = 0;
            for( ; k < rows.size() && rows[k].row == r; ++k )
            {
                for( ; c < rows[k].col; ++c )
                    std::fputs( "0 ", f );
                std::snprintf( num, sizeof( num ), "%g", rows[k].value );
                std::fputs( num, f );
                std::fputc( ' ', f );
                ++c;
            }
            for( ; c < n; ++c )
                std::fputs( "0 ", f );
            std::fputc( '\n', f );
        }
    }
    std::fclose( f );
}
catch( ... )
{
    handle_exception( idx_image, idx_chain );
}

// ===========================================================================
// System.h  (core/src/Spirit/System.cpp)
// ===========================================================================
extern "C" int System_Get_Index( State * state ) noexcept
try
{
    return state->idx_active_image;
}
catch( ... )
{
    handle_exception( -1, -1 );
    return -1;
}

extern "C" int System_Get_NOS( State * state, int idx_image, int idx_chain ) noexcept
try
{
    API_IMAGE();
    return image->nos();
}
catch( ... )
{
    handle_exception( idx_image, idx_chain );
    return 0;
}

extern "C" scalar * System_Get_Spin_Directions( State * state, int idx_image, int idx_chain ) noexcept
try
{
    API_IMAGE();
    sync_image_host( state, idx_image );
    gneb_sync_host( state );
    return image->spins.data();
}
catch( ... )
{
    handle_exception( idx_image, idx_chain );
    return nullptr;
}

extern "C" scalar * System_Get_Effective_Field( State * state, int idx_image, int idx_chain ) noexcept
try
{
    API_IMAGE();
    sync_image_host( state, idx_image );
    gneb_sync_host( state );
    return image->eff_field.data();
}
catch( ... )
{
    handle_exception( idx_image, idx_chain );
    return nullptr;
}

extern "C" scalar * System_Get_Eigenmode( State * state, int idx_mode, int idx_image, int idx_chain ) noexcept
{
    log_message( 2, "System_Get_Eigenmode: eigenmodes are not part of the HIP engine" );
    return nullptr;
}

extern "C" float System_Get_Rx( State * state, int idx_image, int idx_chain ) noexcept
try
{
    API_IMAGE();
    return idx_image < (int)chain->Rx.size() ? (float)chain->Rx[idx_image] : 0.f;
}
catch( ... )
{
    handle_exception( idx_image, idx_chain );
    return 0;
}

extern "C" float System_Get_Energy( State * state, int idx_image, int idx_chain ) noexcept
try
{
    API_IMAGE();
    return (float)image->E;
}
catch( ... )
{
    handle_exception( idx_image, idx_chain );
    return 0;
}

// The contributions present in a Hamiltonian (Hamiltonian_Heisenberg::Update_Energy_Contributions,
// Hamiltonian_Heisenberg.cpp:200-260), as indices into the per-spin terms of local_energy_terms.
std::vector<std::pair<std::string, int>> spirit_hip::energy_contribution_list( const Image & image )
{
    const Hamiltonian & h = *image.hamiltonian;
    std::vector<std::pair<std::string, int>> c;
    if( std::abs( h.field_magnitude ) > 1e-60 )
        c.push_back( { "Zeeman", 0 } );
    if( !h.aniso_idx.empty() || !h.site_K.empty() )
        c.push_back( { "Anisotropy", 1 } );
    if( !h.cubic_idx.empty() )
        c.push_back( { "Cubic anisotropy", 2 } );
    if( !h.exch_pairs.empty() )
        c.push_back( { "Exchange", 3 } );
    if( !h.dmi_pairs.empty() )
        c.push_back( { "DMI", 4 } );
    if( h.ddi_method != DDI_NONE )
        c.push_back( { "DDI", 5 } );
    if( !h.quads.empty() )
        c.push_back( { "Quadruplets", 6 } );
    return c;
}

// The energy of the image's host spins, with the effective field -dE/ds and the energy
// contributions (Spin_System::UpdateEnergy / UpdateEffectiveField, core/src/data/Spin_System.cpp:138-161,
// Hamiltonian_Heisenberg::Energy_Contributions), evaluated on the GPU.
double spirit_hip::update_image_energy( Image & image, bool store_field, std::vector<double> * per_spin )
{
    ImageSpec sp{ image.geometry, image.hamiltonian, *image.llg, image.spins.data() };
    LLGRunner r( { sp }, SOLVER_DEPONDT );
    std::vector<double> g( 3 * (size_t)image.nos() );
    double E = r.gradient( 0, g.data() );
    if( store_field )
        for( size_t k = 0; k < g.size(); ++k )
            image.eff_field[k] = -1 * g[k];
    const size_t nos = (size_t)image.nos();
    std::vector<double> terms( N_ENERGY_TERMS_HOST * nos );
    r.energy_terms( 0, terms.data() );
    image.E_array.clear();
    for( auto & c : energy_contribution_list( image ) )
    {
        double sum = 0;
        for( size_t i = 0; i < nos; ++i )
            sum += terms[(size_t)c.second * nos + i];
        image.E_array.push_back( { c.first, sum } );
    }
    if( per_spin )
        *per_spin = std::move( terms );
    return E;
}

extern "C" int System_Get_Energy_Array_Names( State * state, char * names, int idx_image, int idx_chain ) noexcept
try
{
    API_IMAGE();
    std::string n;
    for( auto & c : energy_contribution_list( *image ) )
        n += ( n.empty() ? "" : "|" ) + c.first;
    if( names )
        std::strcpy( names, n.c_str() );
    return (int)n.size();
}
catch( ... )
{
    handle_exception( idx_image, idx_chain );
    return 0;
}

// System.cpp:179-208: the contributions of the last System_Update_Data (per spin with
// divide_by_nspins); with energies == nullptr the number of contributions
extern "C" int System_Get_Energy_Array( State * state, float * energies, bool divide_by_nspins, int idx_image,
                                        int idx_chain ) noexcept
try
{
    API_IMAGE();
    if( !energies )
        return (int)image->E_array.size();
    const double nd = divide_by_nspins ? 1 / (double)image->nos() : 1;
    for( size_t k = 0; k < image->E_array.size(); ++k )
        energies[k] = (float)( nd * image->E_array[k].second );
    return -1;
}
catch( ... )
{
    handle_exception( idx_image, idx_chain );
    return -1;
}

// Per-spin energy contributions of the image's current spins, [contribution][nos] in the order of
// System_Get_Energy_Array_Names; returns the number of contributions (out may be nullptr).
extern "C" int Spirit_HIP_Get_Energy_Contributions_per_Spin( State * state, double * out, int idx_image,
                                                             int idx_chain ) noexcept
try
{
    API_IMAGE();
    sync_image_host( state, idx_image );
    std::lock_guard<std::recursive_mutex> lk( image->mtx );
    auto list = energy_contribution_list( *image );
    if( !out )
        return (int)list.size();
    std::vector<double> terms;
    image->E = update_image_energy( *image, true, &terms );
    const size_t nos = (size_t)image->nos();
    for( size_t k = 0; k < list.size(); ++k )
        std::memcpy( out + k * nos, terms.data() + (size_t)list[k].second * nos, sizeof( double ) * nos );
    return (int)list.size();
}
catch( ... )
{
    handle_exception( idx_image, idx_chain );
    return 0;
}

extern "C" void System_Get_Eigenvalues( State * state, float * eigenvalues, int idx_image, int idx_chain ) noexcept
{
}

extern "C" void System_Print_Energy_Array( State * state, int idx_image, int idx_chain ) noexcept
try
{
    API_IMAGE();
    // System.cpp:232-256
    const double n = image->nos();
    std::printf( "E_tot = %.12g  ||  E_tot/N = %.12g\n", image->E, image->E / n );
    for( auto & c : image->E_array )
        std::printf( "  %s = %.12g  ||  /N = %.12g\n", c.first.c_str(), c.second, c.second / n );
}
catch( ... )
{
    handle_exception( idx_image, idx_chain );
}

extern "C" void System_Update_Data( State * state, int idx_image, int idx_chain ) noexcept
try
{
    API_IMAGE();
    sync_image_host( state, idx_image );
    std::lock_guard<std::recursive_mutex> lk( image->mtx );
    image->E = update_image_energy( *image, true );
}
catch( ... )
{
    handle_exception( idx_image, idx_chain );
}

extern "C" void System_Update_Eigenmodes( State * state, int idx_image, int idx_chain ) noexcept
{
    log_message( 2, "System_Update_Eigenmodes: eigenmodes are not part of the HIP engine" );
}

// ===========================================================================
// Geometry.h  (core/src/Spirit/Geometry.cpp)
// ===========================================================================
// Rebuild every image after a geometry change (Geometry.cpp helper: new geometry, Hamiltonian
// interactions updated, spins resized to +z where new).
static void apply_geometry( State * state, const Geometry & proto )
{
    auto g = std::make_shared<Geometry>( proto );
    g->generate();
    for( auto & img : state->chain->images )
    {
        std::lock_guard<std::recursive_mutex> lk( img->mtx );
        int old_nos   = img->nos();
        img->geometry = g;
        img->spins.resize( 3 * (size_t)g->nos, 0.0 );
        for( int i = old_nos; i < g->nos; ++i )
            img->spins[3 * i + 2] = 1.0;
        img->eff_field.assign( 3 * (size_t)g->nos, 0.0 );
        auto & h = *img->hamiltonian;
        int nb   = g->nb;
        if( !h.aniso_idx.empty() )
        {
            double K = h.aniso_K[0];
            Vec3 n   = h.aniso_n[0];
            h.aniso_idx.resize( nb );
            h.aniso_K.assign( nb, K );
            h.aniso_n.assign( nb, n );
            for( int i = 0; i < nb; ++i )
                h.aniso_idx[i] = i;
        }
        h.site_K.clear();
        h.update_interactions( *g );
    }
    if( state->clipboard_image )
        state->clipboard_image.reset();
    State_Update( state );
}

static Geometry current_geometry( State * state )
{
    return *state->active_image->geometry;
}

extern "C" void Geometry_Set_Bravais_Lattice_Type( State * state, Bravais_Lattice_Type lattice_type ) noexcept
try
{
    Geometry g = current_geometry( state );
    const double h = 0.5 * std::sqrt( 3.0 );
    switch( lattice_type )
    {
        case Bravais_Lattice_SC:
            g.bravais = { Vec3{ 1, 0, 0 }, Vec3{ 0, 1, 0 }, Vec3{ 0, 0, 1 } };
            g.bravais_type = BravaisType::SC;
            break;
        case Bravais_Lattice_Hex2D:
        case Bravais_Lattice_Hex2D_60:
            g.bravais = { Vec3{ h, -0.5, 0 }, Vec3{ h, 0.5, 0 }, Vec3{ 0, 0, 1 } };
            g.bravais_type = BravaisType::Hex2D;
            break;
        case Bravais_Lattice_Hex2D_120:
            g.bravais = { Vec3{ 0.5, -h, 0 }, Vec3{ 0.5, h, 0 }, Vec3{ 0, 0, 1 } };
            g.bravais_type = BravaisType::Hex2D;
            break;
        case Bravais_Lattice_BCC:
            g.bravais = { Vec3{ 0.5, 0.5, -0.5 }, Vec3{ -0.5, 0.5, -0.5 }, Vec3{ 0.5, -0.5, -0.5 } };
            g.bravais_type = BravaisType::BCC;
            break;
        case Bravais_Lattice_FCC:
            g.bravais = { Vec3{ 0.5, 0.0, 0.5 }, Vec3{ 0.5, 0.5, 0.0 }, Vec3{ 0.0, 0.5, 0.5 } };
            g.bravais_type = BravaisType::FCC;
            break;
        default: throw api_error( 3, "Geometry_Set_Bravais_Lattice_Type: lattice type not available" );
    }
    apply_geometry( state, g );
}
catch( ... )
{
    handle_exception( -1, -1 );
}

extern "C" void Geometry_Set_N_Cells( State * state, int n_cells[3] ) noexcept
try
{
    Geometry g = current_geometry( state );
    for( int c = 0; c < 3; ++c )
    {
        if( n_cells[c] < 1 )
            throw api_error( 2, "Geometry_Set_N_Cells: n_cells must be positive" );
        g.n_cells[c] = n_cells[c];
    }
    apply_geometry( state, g );
}
catch( ... )
{
    handle_exception( -1, -1 );
}

extern "C" void Geometry_Set_Cell_Atoms( State * state, int n_atoms, float ** atoms ) noexcept
try
{
    Geometry g = current_geometry( state );
    if( n_atoms < 1 || n_atoms > MAX_BASIS )
        throw api_error( 2, "Geometry_Set_Cell_Atoms: 1.." + std::to_string( MAX_BASIS ) + " atoms supported" );
    g.cell_atoms.assign( n_atoms, Vec3{ 0, 0, 0 } );
    double mu0 = g.mu_basis.empty() ? 1.0 : g.mu_basis[0];
    g.mu_basis.assign( n_atoms, mu0 );
    for( int a = 0; a < n_atoms; ++a )
        for( int c = 0; c < 3; ++c )
            g.cell_atoms[a][c] = atoms[a][c];
    apply_geometry( state, g );
}
catch( ... )
{
    handle_exception( -1, -1 );
}

extern "C" void Geometry_Set_mu_s( State * state, float mu_s, int idx_image, int idx_chain ) noexcept
try
{
    Geometry g = current_geometry( state );
    for( auto & m : g.mu_basis )
        m = mu_s;
    apply_geometry( state, g );
}
catch( ... )
{
    handle_exception( idx_image, idx_chain );
}

extern "C" void Geometry_Set_Cell_Atom_Types( State * state, int n_atoms, int * atom_types ) noexcept
{
    log_message( 3, "Geometry_Set_Cell_Atom_Types: defects are not enabled in this build" );
}

extern "C" void Geometry_Set_Bravais_Vectors( State * state, float ta[3], float tb[3], float tc[3] ) noexcept
try
{
    Geometry g = current_geometry( state );
    g.bravais  = { Vec3{ ta[0], ta[1], ta[2] }, Vec3{ tb[0], tb[1], tb[2] }, Vec3{ tc[0], tc[1], tc[2] } };
    g.bravais_type = BravaisType::Irregular;
    apply_geometry( state, g );
}
catch( ... )
{
    handle_exception( -1, -1 );
}

extern "C" void Geometry_Set_Lattice_Constant( State * state, float lattice_constant ) noexcept
try
{
    Geometry g         = current_geometry( state );
    g.lattice_constant = lattice_constant;
    apply_geometry( state, g );
}
catch( ... )
{
    handle_exception( -1, -1 );
}

extern "C" int Geometry_Get_NOS( State * state ) noexcept
try
{
    return state->nos;
}
catch( ... )
{
    handle_exception( -1, -1 );
    return 0;
}

extern "C" scalar * Geometry_Get_Positions( State * state, int idx_image, int idx_chain ) noexcept
try
{
    API_IMAGE();
    return image->geometry->positions.data();
}
catch( ... )
{
    handle_exception( idx_image, idx_chain );
    return nullptr;
}

extern "C" int * Geometry_Get_Atom_Types( State * state, int idx_image, int idx_chain ) noexcept
try
{
    API_IMAGE();
    static thread_local std::vector<int> types;
    types.assign( image->nos(), 0 );
    return types.data();
}
catch( ... )
{
    handle_exception( idx_image, idx_chain );
    return nullptr;
}

extern "C" void Geometry_Get_Bounds( State * state, float min[3], float max[3], int idx_image, int idx_chain ) noexcept
try
{
    API_IMAGE();
    for( int c = 0; c < 3; ++c )
    {
        min[c] = (float)image->geometry->bounds_min[c];
        max[c] = (float)image->geometry->bounds_max[c];
    }
}
catch( ... )
{
    handle_exception( idx_image, idx_chain );
}

extern "C" void Geometry_Get_Center( State * state, float center[3], int idx_image, int idx_chain ) noexcept
try
{
    API_IMAGE();
    for( int c = 0; c < 3; ++c )
        center[c] = (float)image->geometry->center[c];
}
catch( ... )
{
    handle_exception( idx_image, idx_chain );
}

extern "C" Bravais_Lattice_Type Geometry_Get_Bravais_Lattice_Type( State * state, int idx_image, int idx_chain ) noexcept
try
{
    API_IMAGE();
    switch( image->geometry->bravais_type )
    {
        case BravaisType::SC: return Bravais_Lattice_SC;
        case BravaisType::Hex2D: return Bravais_Lattice_Hex2D;
        case BravaisType::BCC: return Bravais_Lattice_BCC;
        case BravaisType::FCC: return Bravais_Lattice_FCC;
        default: return Bravais_Lattice_Irregular;
    }
}
catch( ... )
{
    handle_exception( idx_image, idx_chain );
    return Bravais_Lattice_Irregular;
}

extern "C" void Geometry_Get_Bravais_Vectors( State * state, float a[3], float b[3], float c[3], int idx_image,
                                              int idx_chain ) noexcept
try
{
    API_IMAGE();
    for( int q = 0; q < 3; ++q )
    {
        a[q] = (float)image->geometry->bravais[0][q];
        b[q] = (float)image->geometry->bravais[1][q];
        c[q] = (float)image->geometry->bravais[2][q];
    }
}
catch( ... )
{
    handle_exception( idx_image, idx_chain );
}

extern "C" int Geometry_Get_Dimensionality( State * state, int idx_image, int idx_chain ) noexcept
try
{
    API_IMAGE();
    return image->geometry->dimensionality;
}
catch( ... )
{
    handle_exception( idx_image, idx_chain );
    return 0;
}

extern "C" void Geometry_Get_mu_s( State * state, float * mu_s, int idx_image, int idx_chain ) noexcept
try
{
    API_IMAGE();
    for( int b = 0; b < image->geometry->nb; ++b )
        mu_s[b] = (float)image->geometry->mu_basis[b];
}
catch( ... )
{
    handle_exception( idx_image, idx_chain );
}

extern "C" void Geometry_Get_N_Cells( State * state, int n_cells[3], int idx_image, int idx_chain ) noexcept
try
{
    API_IMAGE();
    for( int c = 0; c < 3; ++c )
        n_cells[c] = image->geometry->n_cells[c];
}
catch( ... )
{
    handle_exception( idx_image, idx_chain );
}

extern "C" void Geometry_Get_Cell_Bounds( State * state, float min[3], float max[3], int idx_image,
                                          int idx_chain ) noexcept
try
{
    API_IMAGE();
    for( int c = 0; c < 3; ++c )
    {
        min[c] = (float)image->geometry->cell_bounds_min[c];
        max[c] = (float)image->geometry->cell_bounds_max[c];
    }
}
catch( ... )
{
    handle_exception( idx_image, idx_chain );
}

extern "C" int Geometry_Get_N_Cell_Atoms( State * state, int idx_image, int idx_chain ) noexcept
try
{
    API_IMAGE();
    return image->geometry->nb;
}
catch( ... )
{
    handle_exception( idx_image, idx_chain );
    return 0;
}

extern "C" int Geometry_Get_Cell_Atoms( State * state, scalar ** atoms, int idx_image, int idx_chain ) noexcept
try
{
    API_IMAGE();
    if( atoms )
        *atoms = reinterpret_cast<scalar *>( image->geometry->cell_atoms.data() );
    return image->geometry->nb;
}
catch( ... )
{
    handle_exception( idx_image, idx_chain );
    return 0;
}

// ===========================================================================
// Parameters_LLG.h  (core/src/Spirit/Parameters_LLG.cpp)
// ===========================================================================
#define LLG_SET( body )                                                                                                \
    try                                                                                                                \
    {                                                                                                                  \
        API_IMAGE();                                                                                                   \
        std::lock_guard<std::recursive_mutex> lk( image->mtx );                                                        \
        auto & p = *image->llg;                                                                                        \
        body;                                                                                                          \
        notify_image_changed( state, idx_image, false, false, true );                                                  \
    }                                                                                                                  \
    catch( ... )                                                                                                       \
    {                                                                                                                  \
        handle_exception( idx_image, idx_chain );                                                                      \
    }

extern "C" void Parameters_LLG_Set_Output_Tag( State * state, const char * tag, int idx_image, int idx_chain ) noexcept
{
    LLG_SET( p.output_file_tag = tag );
}
extern "C" void Parameters_LLG_Set_Output_Folder( State * state, const char * folder, int idx_image,
                                                  int idx_chain ) noexcept
{
    LLG_SET( p.output_folder = folder );
}
extern "C" void Parameters_LLG_Set_Output_General( State * state, bool any, bool initial, bool final, int idx_image,
                                                   int idx_chain ) noexcept
{
    LLG_SET( p.output_any = any; p.output_initial = initial; p.output_final = final );
}
extern "C" void Parameters_LLG_Set_Output_Energy( State * state, bool energy_step, bool energy_archive,
                                                  bool energy_spin_resolved, bool energy_divide_by_nos,
                                                  bool energy_add_readability_lines, int idx_image,
                                                  int idx_chain ) noexcept
{
    LLG_SET( p.output_energy_step = energy_step; p.output_energy_archive = energy_archive;
             p.output_energy_spin_resolved = energy_spin_resolved; p.output_energy_divide_by_nspins = energy_divide_by_nos;
             p.output_energy_add_readability_lines = energy_add_readability_lines );
}
extern "C" void Parameters_LLG_Set_Output_Configuration( State * state, bool configuration_step,
                                                         bool configuration_archive, int configuration_filetype,
                                                         int idx_image, int idx_chain ) noexcept
{
    LLG_SET( p.output_configuration_step = configuration_step; p.output_configuration_archive = configuration_archive;
             p.output_configuration_filetype = configuration_filetype );
}
extern "C" void Parameters_LLG_Set_N_Iterations( State * state, int n_iterations, int n_iterations_log, int idx_image,
                                                 int idx_chain ) noexcept
{
    LLG_SET( p.n_iterations = n_iterations; p.n_iterations_log = n_iterations_log );
}
extern "C" void Parameters_LLG_Set_Direct_Minimization( State * state, bool direct, int idx_image,
                                                        int idx_chain ) noexcept
{
    LLG_SET( p.direct_minimization = direct );
}
extern "C" void Parameters_LLG_Set_Convergence( State * state, float convergence, int idx_image, int idx_chain ) noexcept
{
    LLG_SET( p.force_convergence = convergence );
}
extern "C" void Parameters_LLG_Set_Time_Step( State * state, float dt, int idx_image, int idx_chain ) noexcept
{
    LLG_SET( p.dt = dt );
}
extern "C" void Parameters_LLG_Set_Damping( State * state, float damping, int idx_image, int idx_chain ) noexcept
{
    LLG_SET( p.damping = damping );
}
extern "C" void Parameters_LLG_Set_Non_Adiabatic_Damping( State * state, float beta, int idx_image, int idx_chain ) noexcept
{
    LLG_SET( p.beta = beta );
}
extern "C" void Parameters_LLG_Set_STT( State * state, bool use_gradient, float magnitude, const float normal[3],
                                        int idx_image, int idx_chain ) noexcept
{
    LLG_SET( p.stt_use_gradient = use_gradient; p.stt_magnitude = magnitude;
             p.stt_polarisation_normal = stt_normal( normal ) );
}
extern "C" void Parameters_LLG_Set_Temperature( State * state, float T, int idx_image, int idx_chain ) noexcept
{
    LLG_SET( p.temperature = T );
}
extern "C" void Parameters_LLG_Set_Temperature_Gradient( State * state, float inclination, const float direction[3],
                                                         int idx_image, int idx_chain ) noexcept
{
    // stored as given; the distribution normalises it (Vectormath.cpp:638)
    LLG_SET( p.temperature_gradient_inclination = inclination;
             p.temperature_gradient_direction   = as_vec3( direction ) );
}

#define LLG_GET( ret, expr )                                                                                           \
    try                                                                                                                \
    {                                                                                                                  \
        API_IMAGE();                                                                                                   \
        auto & p = *image->llg;                                                                                        \
        return expr;                                                                                                   \
    }                                                                                                                  \
    catch( ... )                                                                                                       \
    {                                                                                                                  \
        handle_exception( idx_image, idx_chain );                                                                      \
        return ret;                                                                                                    \
    }

extern "C" const char * Parameters_LLG_Get_Output_Tag( State * state, int idx_image, int idx_chain ) noexcept
{
    LLG_GET( "", p.output_file_tag.c_str() );
}
extern "C" const char * Parameters_LLG_Get_Output_Folder( State * state, int idx_image, int idx_chain ) noexcept
{
    LLG_GET( "", p.output_folder.c_str() );
}
extern "C" void Parameters_LLG_Get_Output_Energy( State * state, bool * energy_step, bool * energy_archive,
                                                  bool * energy_spin_resolved, bool * energy_divide_by_nos,
                                                  bool * energy_add_readability_lines, int idx_image,
                                                  int idx_chain ) noexcept
try
{
    API_IMAGE();
    const auto & p                = *image->llg;
    *energy_step                  = p.output_energy_step;
    *energy_archive               = p.output_energy_archive;
    *energy_spin_resolved         = p.output_energy_spin_resolved;
    *energy_divide_by_nos         = p.output_energy_divide_by_nspins;
    *energy_add_readability_lines = p.output_energy_add_readability_lines;
}
catch( ... )
{
    handle_exception( idx_image, idx_chain );
}
extern "C" void Parameters_LLG_Get_Output_Configuration( State * state, bool * configuration_step,
                                                         bool * configuration_archive, int * configuration_filetype,
                                                         int idx_image, int idx_chain ) noexcept
try
{
    API_IMAGE();
    *configuration_step     = image->llg->output_configuration_step;
    *configuration_archive  = image->llg->output_configuration_archive;
    *configuration_filetype = image->llg->output_configuration_filetype;
}
catch( ... )
{
    handle_exception( idx_image, idx_chain );
}
extern "C" void Parameters_LLG_Get_Output_General( State * state, bool * any, bool * initial, bool * final,
                                                   int idx_image, int idx_chain ) noexcept
try
{
    API_IMAGE();
    *any     = image->llg->output_any;
    *initial = image->llg->output_initial;
    *final   = image->llg->output_final;
}
catch( ... )
{
    handle_exception( idx_image, idx_chain );
}
extern "C" void Parameters_LLG_Get_N_Iterations( State * state, int * iterations, int * iterations_log, int idx_image,
                                                 int idx_chain ) noexcept
try
{
    API_IMAGE();
    *iterations     = (int)image->llg->n_iterations;
    *iterations_log = (int)image->llg->n_iterations_log;
}
catch( ... )
{
    handle_exception( idx_image, idx_chain );
}
extern "C" bool Parameters_LLG_Get_Direct_Minimization( State * state, int idx_image, int idx_chain ) noexcept
{
    LLG_GET( false, p.direct_minimization );
}
extern "C" float Parameters_LLG_Get_Convergence( State * state, int idx_image, int idx_chain ) noexcept
{
    LLG_GET( 0.f, (float)p.force_convergence );
}
extern "C" float Parameters_LLG_Get_Time_Step( State * state, int idx_image, int idx_chain ) noexcept
{
    LLG_GET( 0.f, (float)p.dt );
}
extern "C" float Parameters_LLG_Get_Damping( State * state, int idx_image, int idx_chain ) noexcept
{
    LLG_GET( 0.f, (float)p.damping );
}
extern "C" float Parameters_LLG_Get_Non_Adiabatic_Damping( State * state, int idx_image, int idx_chain ) noexcept
{
    LLG_GET( 0.f, (float)p.beta );
}
extern "C" float Parameters_LLG_Get_Temperature( State * state, int idx_image, int idx_chain ) noexcept
{
    LLG_GET( 0.f, (float)p.temperature );
}
extern "C" void Parameters_LLG_Get_Temperature_Gradient( State * state, float * inclination, float direction[3],
                                                         int idx_image, int idx_chain ) noexcept
try
{
    API_IMAGE();
    *inclination = (float)image->llg->temperature_gradient_inclination;
    for( int c = 0; c < 3; ++c )
        direction[c] = (float)image->llg->temperature_gradient_direction[c];
}
catch( ... )
{
    handle_exception( idx_image, idx_chain );
}
extern "C" void Parameters_LLG_Get_STT( State * state, bool * use_gradient, float * magnitude, float normal[3],
                                        int idx_image, int idx_chain ) noexcept
try
{
    API_IMAGE();
    *use_gradient = image->llg->stt_use_gradient;
    *magnitude    = (float)image->llg->stt_magnitude;
    for( int c = 0; c < 3; ++c )
        normal[c] = (float)image->llg->stt_polarisation_normal[c];
}
catch( ... )
{
    handle_exception( idx_image, idx_chain );
}

// ===========================================================================
// Quantities.h  (core/src/Spirit/Quantities.cpp, Vectormath::Magnetization)
// ===========================================================================
extern "C" void Quantity_Get_Average_Spin( State * state, float s[3], int idx_image, int idx_chain ) noexcept
try
{
    API_IMAGE();
    sync_image_host( state, idx_image );
    gneb_sync_host( state );
    double m[3] = { 0, 0, 0 };
    for( int i = 0; i < image->nos(); ++i )
        for( int c = 0; c < 3; ++c )
            m[c] += image->spins[3 * i + c];
    for( int c = 0; c < 3; ++c )
        s[c] = (float)( m[c] / image->nos() );
}
catch( ... )
{
    handle_exception( idx_image, idx_chain );
}

extern "C" void Quantity_Get_Magnetization( State * state, float m[3], int idx_image, int idx_chain ) noexcept
try
{
    API_IMAGE();
    sync_image_host( state, idx_image );
    gneb_sync_host( state );
    // Vectormath::Magnetization: mean of mu_s * s
    double M[3] = { 0, 0, 0 };
    for( int i = 0; i < image->nos(); ++i )
        for( int c = 0; c < 3; ++c )
            M[c] += image->geometry->mu_s[i] * image->spins[3 * i + c];
    for( int c = 0; c < 3; ++c )
        m[c] = (float)( M[c] / image->nos() );
}
catch( ... )
{
    handle_exception( idx_image, idx_chain );
}

// Topological charge density of a 2D lattice (one basis atom): per triangle
// (a,b)-(a+1,b)-(a,b+1) and (a+1,b)-(a+1,b+1)-(a,b+1), its solid angle (Oosterom-Strackee,
// Vectormath.cpp solid_angle_2) over 4 pi, oriented so that a counter-clockwise triangle in the
// xy plane counts positive; periodic directions wrap.  Returns the number of triangles; the
// densities and the triangles' spin indices (3 per triangle) go to the non-null outputs.
static int topological_density( Image & image, std::vector<double> * density, std::vector<int> * triangles )
{
    auto & g = *image.geometry;
    auto & h = *image.hamiltonian;
    if( g.nb != 1 || g.n_cells[2] != 1 )
        return 0;
    const int Na = g.n_cells[0], Nb = g.n_cells[1];
    auto S = [&]( int idx ) -> const double * { return &image.spins[3 * (size_t)idx]; };
    auto solid = []( const double * v1, const double * v2, const double * v3 ) {
        double c[3] = { v2[1] * v3[2] - v2[2] * v3[1], v2[2] * v3[0] - v2[0] * v3[2], v2[0] * v3[1] - v2[1] * v3[0] };
        double x    = v1[0] * c[0] + v1[1] * c[1] + v1[2] * c[2];
        double y    = 1 + ( v1[0] * v2[0] + v1[1] * v2[1] + v1[2] * v2[2] )
                   + ( v1[0] * v3[0] + v1[1] * v3[1] + v1[2] * v3[2] )
                   + ( v2[0] * v3[0] + v2[1] * v3[1] + v2[2] * v3[2] );
        return 2 * std::atan2( x, y );
    };
    const int amax = h.bc[0] ? Na : Na - 1, bmax = h.bc[1] ? Nb : Nb - 1;
    // orientation: the triangle (a,b),(a+1,b),(a,b+1) is counter-clockwise when a x b points along +z
    const double cz   = g.bravais[0][0] * g.bravais[1][1] - g.bravais[0][1] * g.bravais[1][0];
    const double sign = cz >= 0 ? 1.0 : -1.0;
    int n             = 0;
    for( int b = 0; b < bmax; ++b )
        for( int a = 0; a < amax; ++a )
        {
            const int a1 = ( a + 1 ) % Na, b1 = ( b + 1 ) % Nb;
            const int tri[2][3] = { { a + Na * b, a1 + Na * b, a + Na * b1 }, { a1 + Na * b, a1 + Na * b1, a + Na * b1 } };
            for( auto & t : tri )
            {
                if( density )
                    density->push_back( sign * solid( S( t[0] ), S( t[1] ), S( t[2] ) ) / ( 4 * Pi ) );
                if( triangles )
                    triangles->insert( triangles->end(), t, t + 3 );
                ++n;
            }
        }
    return n;
}

extern "C" float Quantity_Get_Topological_Charge( State * state, int idx_image, int idx_chain ) noexcept
try
{
    API_IMAGE();
    sync_image_host( state, idx_image );
    gneb_sync_host( state );
    std::vector<double> density;
    topological_density( *image, &density, nullptr );
    double q = 0;
    for( double d : density )
        q += d;
    return (float)q;
}
catch( ... )
{
    handle_exception( idx_image, idx_chain );
    return 0;
}

// Quantities.h: the number of triangles; with non-null arrays, each triangle's charge and its
// three spin indices
extern "C" int Quantity_Get_Topological_Charge_Density( State * state, float * charge_density,
                                                        int * triangle_indices, int idx_image, int idx_chain ) noexcept
try
{
    API_IMAGE();
    sync_image_host( state, idx_image );
    gneb_sync_host( state );
    std::vector<double> density;
    std::vector<int> tri;
    const int n = topological_density( *image, &density, &tri );
    if( charge_density )
        for( int k = 0; k < n; ++k )
            charge_density[k] = (float)density[k];
    if( triangle_indices )
        std::copy( tri.begin(), tri.end(), triangle_indices );
    return n;
}
catch( ... )
{
    handle_exception( idx_image, idx_chain );
    return 0;
}

// ===========================================================================
// Constants.h, Version.h, Log.h
// ===========================================================================
extern "C" scalar Constants_mu_B() noexcept
{
    return mu_B;
}
extern "C" scalar Constants_mu_0() noexcept
{
    return mu_0;
}
extern "C" scalar Constants_k_B() noexcept
{
    return k_B;
}
extern "C" scalar Constants_hbar() noexcept
{
    return 0.6582119514;
}
extern "C" scalar Constants_mRy() noexcept
{
    return 1.0 / 13.605693009;
}
extern "C" scalar Constants_gamma() noexcept
{
    return gamma_e;
}
extern "C" scalar Constants_g_e() noexcept
{
    return 2.00231930436182;
}
extern "C" scalar Constants_Pi() noexcept
{
    return Pi;
}

extern "C" int Spirit_Version_Major() noexcept
{
    return 2;
}
extern "C" int Spirit_Version_Minor() noexcept
{
    return 2;
}
extern "C" int Spirit_Version_Patch() noexcept
{
    return 0;
}
extern "C" const char * Spirit_Version() noexcept
{
    return "2.2.0";
}
extern "C" const char * Spirit_Version_Revision() noexcept
{
    return "hip";
}
extern "C" const char * Spirit_Version_Full() noexcept
{
    return "2.2.0-hip";
}
extern "C" const char * Spirit_Compiler() noexcept
{
    return "hipcc";
}
extern "C" const char * Spirit_Compiler_Version() noexcept
{
    return __VERSION__;
}
extern "C" const char * Spirit_Compiler_Full() noexcept
{
    return "hipcc " __VERSION__;
}
extern "C" const char * Spirit_Scalar_Type() noexcept
{
    return "double";
}
extern "C" const char * Spirit_Defects() noexcept
{
    return "OFF";
}
extern "C" const char * Spirit_Pinning() noexcept
{
    return "OFF";
}
extern "C" const char * Spirit_Cuda() noexcept
{
    return "OFF";
}
extern "C" const char * Spirit_OpenMP() noexcept
{
    return "OFF";
}
extern "C" int Spirit_OpenMP_Get_Num_Threads() noexcept
{
    return 1;
}
extern "C" const char * Spirit_Threads() noexcept
{
    return "OFF";
}
extern "C" const char * Spirit_FFTW() noexcept
{
    return "OFF";
}

extern "C" void Log_Send( State * state, Spirit_Log_Level level, Spirit_Log_Sender sender, const char * message,
                          int idx_image, int idx_chain ) noexcept
{
    log_message( (int)level, message ? message : "" );
}
extern "C" void Log_Append( State * state ) noexcept {}
extern "C" void Log_Dump( State * state ) noexcept {}
extern "C" int Log_Get_N_Entries( State * state ) noexcept
{
    return g_n_errors + g_n_warnings;
}
extern "C" int Log_Get_N_Errors( State * state ) noexcept
{
    return g_n_errors;
}
extern "C" int Log_Get_N_Warnings( State * state ) noexcept
{
    return g_n_warnings;
}
extern "C" void Log_Set_Output_To_Console( State * state, bool output, int level ) noexcept
{
    g_log_console_level = output ? level : -1;
}
extern "C" bool Log_Get_Output_To_Console( State * state ) noexcept
{
    return g_log_console_level >= 0;
}
extern "C" int Log_Get_Output_Console_Level( State * state ) noexcept
{
    return g_log_console_level;
}

// ===========================================================================
// HIP-only extensions (include/Spirit/Spirit_HIP.h)
// ===========================================================================
extern "C" void Spirit_HIP_Set_Host_Sync( State * state, int mode ) noexcept
try
{
    state->host_sync = mode ? 1 : 0;
    if( mode )
        sync_all_hosts( state );
}
catch( ... )
{
    handle_exception( -1, -1 );
}

extern "C" void Spirit_HIP_Set_Thermal_Noise( State * state, const double * noise, int n_steps, int idx_image,
                                              int idx_chain ) noexcept
try
{
    API_IMAGE();
    auto it = state->method_image.find( idx_image );
    if( it == state->method_image.end() || !it->second || !it->second->runner )
        throw api_error( 2, "Spirit_HIP_Set_Thermal_Noise: start a single-shot LLG simulation first" );
    it->second->runner->set_injected_noise( noise, n_steps );
}
catch( ... )
{
    handle_exception( idx_image, idx_chain );
}

extern "C" void Spirit_HIP_Hamiltonian_Set_Site_Anisotropy( State * state, const double * K, const float normal[3],
                                                            int idx_image, int idx_chain ) noexcept
try
{
    API_IMAGE();
    std::lock_guard<std::recursive_mutex> lk( image->mtx );
    auto & h = *image->hamiltonian;
    if( K )
    {
        h.site_K.assign( K, K + image->nos() );
        h.site_n = normalized3( normal );
    }
    else
        h.site_K.clear();
    ++h.version;
    notify_image_changed( state, idx_image, false, true, false );
}
catch( ... )
{
    handle_exception( idx_image, idx_chain );
}

extern "C" int Spirit_HIP_Gradient_and_Energy( State * state, double * gradient, double * energy, int idx_image,
                                               int idx_chain ) noexcept
try
{
    API_IMAGE();
    sync_image_host( state, idx_image );
    ImageSpec sp{ image->geometry, image->hamiltonian, *image->llg, image->spins.data() };
    LLGRunner r( { sp }, SOLVER_DEPONDT );
    double E = r.gradient( 0, gradient );
    if( energy )
        *energy = E;
    return 0;
}
catch( ... )
{
    handle_exception( idx_image, idx_chain );
    return 1;
}

extern "C" int Spirit_HIP_Hamiltonian_Get_Hessian( State * state, double * hessian, int idx_image,
                                                   int idx_chain ) noexcept
try
{
    API_IMAGE();
    if( !hessian )
        throw api_error( 2, "Spirit_HIP_Hamiltonian_Get_Hessian: no output array" );
    std::lock_guard<std::recursive_mutex> lk( image->mtx );
    dense_hessian( *image->geometry, *image->hamiltonian, hessian );
    return 0;
}
catch( ... )
{
    handle_exception( idx_image, idx_chain );
    return 1;
}

extern "C" int Spirit_HIP_Hamiltonian_Hessian_Vector_Product( State * state, const double * v, double * out,
                                                              int idx_image, int idx_chain ) noexcept
try
{
    API_IMAGE();
    if( !v || !out )
        throw api_error( 2, "Spirit_HIP_Hamiltonian_Hessian_Vector_Product: no input or output array" );
    std::lock_guard<std::recursive_mutex> lk( image->mtx );
    hessian_vector_product( *image->geometry, *image->hamiltonian, v, out );
    return 0;
}
catch( ... )
{
    handle_exception( idx_image, idx_chain );
    return 1;
}

extern "C" double Spirit_HIP_Get_Last_Step_Time( State * state, int idx_image, int idx_chain ) noexcept
try
{
    API_IMAGE();
    auto it = state->method_image.find( idx_image );
    if( it != state->method_image.end() && it->second && it->second->runner )
        return it->second->runner->last_step_ms;
    return 0;
}
catch( ... )
{
    handle_exception( idx_image, idx_chain );
    return 0;
}

static LLGRunner * runner_of( State * state, int idx_image )
{
    auto it = state->method_image.find( idx_image );
    if( it == state->method_image.end() || !it->second || !it->second->runner )
        throw api_error( 2, "no running LLG method on image " + std::to_string( idx_image ) );
    return it->second->runner.get();
}

extern "C" void Spirit_HIP_Profile_Enable( State * state, bool enable, int idx_image, int idx_chain ) noexcept
try
{
    API_IMAGE();
    runner_of( state, idx_image )->set_profiling( enable );
}
catch( ... )
{
    handle_exception( idx_image, idx_chain );
}

extern "C" void Spirit_HIP_Profile_Get( State * state, const char * kernel, double * total_ms, int * count,
                                        int idx_image, int idx_chain ) noexcept
try
{
    API_IMAGE();
    auto e    = runner_of( state, idx_image )->profile( kernel ? kernel : "" );
    *total_ms = e.ms;
    *count    = (int)e.count;
}
catch( ... )
{
    handle_exception( idx_image, idx_chain );
}

extern "C" void Spirit_HIP_Profile_Reset( State * state, int idx_image, int idx_chain ) noexcept
try
{
    API_IMAGE();
    runner_of( state, idx_image )->reset_profile();
}
catch( ... )
{
    handle_exception( idx_image, idx_chain );
}

extern "C" int Spirit_HIP_DDI_Info( State * state, char * buffer, int length, int idx_image, int idx_chain ) noexcept
try
{
    API_IMAGE();
    const DdiPlan * plan = runner_of( state, idx_image )->ddi_plan( 0 );
    if( !plan )
        return 0;
    const DdiDev & d = plan->dev();
    char msg[256];
    if( plan->real_space() )
    {
        const DdiRsDev & r = plan->rs_dev();
        if( r.method == DDI_RS_DIRECT )
            std::snprintf( msg, sizeof( msg ), "direct sum over %d spins; periodic images %d %d %d", r.nos, r.img[0],
                           r.img[1], r.img[2] );
        else
            std::snprintf( msg, sizeof( msg ), "cutoff: %d pairs", r.n_pairs );
    }
    else
        std::snprintf( msg, sizeof( msg ),
                       "padded %dx%dx%d; spectrum %s; stored %dx%dx%d; pass C kernel %d tile %d (%.3f ms)", d.Pa, d.Pb,
                       d.Pc, plan->spectrum_format(), d.KB, d.Ha, d.KC, d.c_kind, d.c_tk, plan->pass_c_ms() );
    if( buffer && length > 0 )
    {
        std::strncpy( buffer, msg, (size_t)length - 1 );
        buffer[length - 1] = 0;
    }
    return 1;
}
catch( ... )
{
    handle_exception( idx_image, idx_chain );
    return 0;
}

extern "C" const char * Spirit_HIP_Device_Name() noexcept
{
    static std::string name;
    int n = 0;
    if( hipGetDeviceCount( &n ) != hipSuccess || n == 0 )
        return "";
    hipDeviceProp_t prop;
    if( hipGetDeviceProperties( &prop, 0 ) != hipSuccess )
        return "";
    name = std::string( prop.name ) + " (" + prop.gcnArchName + ")";
    return name.c_str();
}

// ---------------------------------------------------------------------------
// Parameters_MC (core/src/Spirit/Parameters_MC.cpp)
// ---------------------------------------------------------------------------
#include "../../include/Spirit/Parameters_MC.h"

extern "C" void Parameters_MC_Set_N_Iterations( State * state, int n_iterations, int n_iterations_log, int idx_image,
                                                int idx_chain ) noexcept
try
{
    API_IMAGE();
    std::lock_guard<std::recursive_mutex> lk( image->mtx );
    image->mc->n_iterations     = n_iterations;
    image->mc->n_iterations_log = n_iterations_log;
}
catch( ... )
{
    handle_exception( idx_image, idx_chain );
}

extern "C" void Parameters_MC_Set_Temperature( State * state, float T, int idx_image, int idx_chain ) noexcept
try
{
    API_IMAGE();
    std::lock_guard<std::recursive_mutex> lk( image->mtx );
    image->mc->temperature = T;
}
catch( ... )
{
    handle_exception( idx_image, idx_chain );
}

extern "C" void Parameters_MC_Set_Metropolis_Cone( State * state, bool cone, float cone_angle, bool adaptive_cone,
                                                   float target_acceptance_ratio, int idx_image,
                                                   int idx_chain ) noexcept
try
{
    API_IMAGE();
    std::lock_guard<std::recursive_mutex> lk( image->mtx );
    image->mc->metropolis_step_cone     = cone;
    image->mc->metropolis_cone_angle    = cone_angle;
    image->mc->metropolis_cone_adaptive = adaptive_cone;
    image->mc->acceptance_ratio_target  = target_acceptance_ratio;
}
catch( ... )
{
    handle_exception( idx_image, idx_chain );
}

extern "C" void Parameters_MC_Set_Random_Sample( State * state, bool random_sample, int idx_image,
                                                 int idx_chain ) noexcept
try
{
    API_IMAGE();
    std::lock_guard<std::recursive_mutex> lk( image->mtx );
    image->mc->metropolis_random_sample = random_sample;
}
catch( ... )
{
    handle_exception( idx_image, idx_chain );
}

extern "C" void Parameters_MC_Get_N_Iterations( State * state, int * iterations, int * iterations_log, int idx_image,
                                                int idx_chain ) noexcept
try
{
    API_IMAGE();
    *iterations     = (int)image->mc->n_iterations;
    *iterations_log = (int)image->mc->n_iterations_log;
}
catch( ... )
{
    handle_exception( idx_image, idx_chain );
}

extern "C" float Parameters_MC_Get_Temperature( State * state, int idx_image, int idx_chain ) noexcept
try
{
    API_IMAGE();
    return (float)image->mc->temperature;
}
catch( ... )
{
    handle_exception( idx_image, idx_chain );
    return 0;
}

extern "C" void Parameters_MC_Get_Metropolis_Cone( State * state, bool * cone, float * cone_angle,
                                                   bool * adaptive_cone, float * target_acceptance_ratio,
                                                   int idx_image, int idx_chain ) noexcept
try
{
    API_IMAGE();
    *cone                    = image->mc->metropolis_step_cone;
    *cone_angle              = (float)image->mc->metropolis_cone_angle;
    *adaptive_cone           = image->mc->metropolis_cone_adaptive;
    *target_acceptance_ratio = (float)image->mc->acceptance_ratio_target;
}
catch( ... )
{
    handle_exception( idx_image, idx_chain );
}

extern "C" bool Parameters_MC_Get_Random_Sample( State * state, int idx_image, int idx_chain ) noexcept
try
{
    API_IMAGE();
    return image->mc->metropolis_random_sample;
}
catch( ... )
{
    handle_exception( idx_image, idx_chain );
    return false;
}
