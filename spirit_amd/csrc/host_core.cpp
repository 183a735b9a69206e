// Host-side geometry, neighbour tables, Hamiltonian parameters and input-file reader.
#include "host_core.hpp"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <limits>
#include <cstdio>
#include <fstream>
#include <random>

namespace spirit_hip
{

int g_log_console_level = 5;
bool g_quiet            = false;
int g_n_errors          = 0;
int g_n_warnings        = 0;

// Log levels follow core/include/utility/Logging.hpp: All=0, Severe=1, Error=2, Warning=3,
// Parameter=4, Info=5, Debug=6.
void log_message( int level, const std::string & msg )
{
    if( level == 2 || level == 1 )
        ++g_n_errors;
    if( level == 3 )
        ++g_n_warnings;
    if( g_quiet && level > 2 )
        return;
    if( level <= g_log_console_level )
        std::fprintf( stderr, "[spirit-hip] %s\n", msg.c_str() );
}

// ---------------------------------------------------------------------------
// Geometry
// ---------------------------------------------------------------------------
void Geometry::basis_position( int iatom, double * out ) const
{
    for( int c = 0; c < 3; ++c )
    {
        double v = ( 0.0 + cell_atoms[iatom][0] ) * bravais[0][c] + ( 0.0 + cell_atoms[iatom][1] ) * bravais[1][c]
                   + ( 0.0 + cell_atoms[iatom][2] ) * bravais[2][c];
        out[c] = lattice_constant * v;
    }
}

// positions: Geometry.cpp:142-161
void Geometry::generate()
{
    nb            = (int)cell_atoms.size();
    n_cells_total = n_cells[0] * n_cells[1] * n_cells[2];
    nos           = nb * n_cells_total;
    positions.assign( 3 * (size_t)nos, 0.0 );
    mu_s.assign( (size_t)nos, 1.0 );
    if( (int)mu_basis.size() < nb )
        mu_basis.resize( nb, mu_basis.empty() ? 1.0 : mu_basis[0] );
    for( long dc = 0; dc < n_cells[2]; ++dc )
        for( long db = 0; db < n_cells[1]; ++db )
            for( long da = 0; da < n_cells[0]; ++da )
                for( int ia = 0; ia < nb; ++ia )
                {
                    long ispin = ia + dc * nb * n_cells[1] * n_cells[0] + db * nb * n_cells[0] + da * nb;
                    for( int c = 0; c < 3; ++c )
                    {
                        double v = ( (double)da + cell_atoms[ia][0] ) * bravais[0][c]
                                   + ( (double)db + cell_atoms[ia][1] ) * bravais[1][c]
                                   + ( (double)dc + cell_atoms[ia][2] ) * bravais[2][c];
                        positions[3 * ispin + c] = lattice_constant * v;
                    }
                    mu_s[ispin] = mu_basis[ia];
                }
    // Geometry::calculateBounds (Geometry.cpp:729-744): the bounds start at the origin
    for( int c = 0; c < 3; ++c )
    {
        bounds_min[c] = 0;
        bounds_max[c] = 0;
    }
    for( int i = 0; i < nos; ++i )
        for( int c = 0; c < 3; ++c )
        {
            bounds_min[c] = std::min( bounds_min[c], positions[3 * i + c] );
            bounds_max[c] = std::max( bounds_max[c], positions[3 * i + c] );
        }
    for( int c = 0; c < 3; ++c )
    {
        center[c]          = 0.5 * ( bounds_min[c] + bounds_max[c] );
        cell_bounds_min[c] = 0;
        cell_bounds_max[c] = 0;
    }
    // Geometry::calculateUnitCellBounds (Geometry.cpp:745-769): half the extent of the basis
    // atoms shifted by +-a_k (and the origin)
    for( int k = 0; k < 3; ++k )
        for( int ia = 0; ia < nb; ++ia )
        {
            double p[3];
            basis_position( ia, p );
            for( int c = 0; c < 3; ++c )
            {
                const double t  = lattice_constant * bravais[k][c];
                const double n1 = p[c] + t, n2 = p[c] - t;
                cell_bounds_min[c] = std::min( cell_bounds_min[c], std::min( n1, n2 ) );
                cell_bounds_max[c] = std::max( cell_bounds_max[c], std::max( n1, n2 ) );
            }
        }
    for( int c = 0; c < 3; ++c )
    {
        cell_bounds_min[c] *= 0.5;
        cell_bounds_max[c] *= 0.5;
    }
    dimensionality = 0;
    for( int c = 0; c < 3; ++c )
        if( n_cells[c] > 1 )
            ++dimensionality;
}

// ---------------------------------------------------------------------------
// Neighbours, core/src/engine/Neighbours.cpp:15-156
// ---------------------------------------------------------------------------
static double norm3( const double * v )
{
    return std::sqrt( v[0] * v[0] + v[1] * v[1] + v[2] * v[2] );
}

std::vector<double> shell_radii( const Geometry & g, int n_shells )
{
    const double min_shell_width = 1e-3;
    std::vector<double> radii( n_shells, 0.0 );
    double ta[3], tb[3], tc[3];
    for( int c = 0; c < 3; ++c )
    {
        ta[c] = g.lattice_constant * g.bravais[0][c];
        tb[c] = g.lattice_constant * g.bravais[1][c];
        tc[c] = g.lattice_constant * g.bravais[2][c];
    }
    int max_n = n_shells + 2;
    int i_max = std::min( max_n, g.n_cells[0] - 1 );
    int j_max = std::min( max_n, g.n_cells[1] - 1 );
    int k_max = std::min( max_n, g.n_cells[2] - 1 );
    if( norm3( ta ) == 0.0 )
        i_max = 0;
    if( norm3( tb ) == 0.0 )
        j_max = 0;
    if( norm3( tc ) == 0.0 )
        k_max = 0;
    double outermost = 0, previous = 0;
    for( auto & radius : radii )
    {
        previous  = outermost;
        outermost = DBL_MAX;
        for( int a1 = 0; a1 < (int)g.cell_atoms.size(); ++a1 )
        {
            const Vec3 & pos_one = g.cell_atoms[a1];
            for( int i = i_max; i >= 0; --i )
                for( int j = j_max; j >= -j_max; --j )
                    for( int k = k_max; k >= -k_max; --k )
                        for( int a2 = 0; a2 < (int)g.cell_atoms.size(); ++a2 )
                        {
                            if( a1 == a2 && i == 0 && j == 0 && k == 0 )
                                continue;
                            double d[3];
                            for( int c = 0; c < 3; ++c )
                            {
                                double pos_two = g.cell_atoms[a2][c] + i * ta[c] + j * tb[c] + k * tc[c];
                                d[c]           = pos_one[c] - pos_two;
                            }
                            double delta = norm3( d );
                            if( delta - previous > min_shell_width && delta < outermost )
                            {
                                outermost = delta;
                                radius    = delta;
                            }
                        }
        }
    }
    return radii;
}

void neighbours_in_shells( const Geometry & g, int n_shells, bool redundant, std::vector<Pair> & pairs,
                           std::vector<int> & shells )
{
    const double min_shell_width = 1e-3;
    auto radii                   = shell_radii( g, n_shells );
    double ta[3], tb[3], tc[3];
    for( int c = 0; c < 3; ++c )
    {
        ta[c] = g.lattice_constant * g.bravais[0][c];
        tb[c] = g.lattice_constant * g.bravais[1][c];
        tc[c] = g.lattice_constant * g.bravais[2][c];
    }
    int max_n = n_shells + 2;
    int i_max = std::min( max_n, g.n_cells[0] - 1 );
    int j_max = std::min( max_n, g.n_cells[1] - 1 );
    int k_max = std::min( max_n, g.n_cells[2] - 1 );
    int i_min = -i_max, j_min = -j_max, k_min = -k_max;
    if( norm3( ta ) == 0.0 )
        i_max = 0;
    if( norm3( tb ) == 0.0 )
        j_max = 0;
    if( norm3( tc ) == 0.0 )
        k_max = 0;
    const int nb = (int)g.cell_atoms.size();
    for( int a1 = 0; a1 < nb; ++a1 )
    {
        int second_atom_min = redundant ? 0 : a1;
        const Vec3 & pos_one = g.cell_atoms[a1];
        for( int s = 0; s < n_shells; ++s )
        {
            double radius = radii[s];
            for( int i = i_max; i >= i_min; --i )
                for( int j = j_max; j >= j_min; --j )
                    for( int k = k_max; k >= k_min; --k )
                        for( int a2 = second_atom_min; a2 < nb; ++a2 )
                        {
                            if( ( a2 > a1 ) || ( i > 0 || ( i == 0 && j > 0 ) || ( i == 0 && j == 0 && k > 0 ) )
                                || redundant )
                            {
                                double d[3];
                                for( int c = 0; c < 3; ++c )
                                {
                                    double pos_two = g.cell_atoms[a2][c] + i * ta[c] + j * tb[c] + k * tc[c];
                                    d[c]           = pos_one[c] - pos_two;
                                }
                                double delta = norm3( d );
                                if( std::fabs( delta - radius ) < min_shell_width )
                                {
                                    pairs.push_back( Pair{ a1, a2, { i, j, k } } );
                                    shells.push_back( s );
                                }
                            }
                        }
        }
    }
}

// DMI_Normal_from_Pair, Neighbours.cpp:249-286
Vec3 dmi_normal_from_pair( const Geometry & g, const Pair & p, int chirality )
{
    double ipos[3], jpos[3], d[3];
    g.basis_position( p.i, ipos );
    g.basis_position( p.j, jpos );
    for( int c = 0; c < 3; ++c )
    {
        double ta = g.lattice_constant * g.bravais[0][c];
        double tb = g.lattice_constant * g.bravais[1][c];
        double tc = g.lattice_constant * g.bravais[2][c];
        jpos[c]   = jpos[c] + p.t[0] * ta + p.t[1] * tb + p.t[2] * tc;
    }
    auto normalized = []( double * v ) {
        double z = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
        if( z > 0 )
        {
            double n = std::sqrt( z );
            v[0] /= n;
            v[1] /= n;
            v[2] /= n;
        }
    };
    auto cross = []( const double * a, const double * b ) {
        return Vec3{ a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0] };
    };
    const double ez[3] = { 0, 0, 1 };
    if( chirality == 1 || chirality == 2 || chirality == -2 )
    {
        for( int c = 0; c < 3; ++c )
            d[c] = jpos[c] - ipos[c];
        normalized( d );
        if( chirality == 1 )
            return Vec3{ d[0], d[1], d[2] };
        if( chirality == 2 )
            return cross( d, ez );
        return cross( ez, d );
    }
    if( chirality == -1 )
    {
        for( int c = 0; c < 3; ++c )
            d[c] = ipos[c] - jpos[c];
        normalized( d );
        return Vec3{ d[0], d[1], d[2] };
    }
    return Vec3{ 0, 0, 0 };
}

// Hamiltonian_Heisenberg::Update_Interactions, Hamiltonian_Heisenberg.cpp:101-198
void pairs_in_radius( const Geometry & g, double radius, std::vector<Pair> & pairs, std::vector<double> & dist,
                      std::vector<Vec3> & normal )
{
    pairs.clear();
    dist.clear();
    normal.clear();
    if( std::abs( radius ) < 1e-6 )
    {
        log_message( 0, "Generating pairs within a radius of less than 1e-06 is not supported" );
        return;
    }
    // translation vectors and how many of them can reach the radius (from the lattice extent)
    Vec3 tv[3];
    int span[3] = { 0, 0, 0 };
    for( int k = 0; k < 3; ++k )
    {
        for( int c = 0; c < 3; ++c )
            tv[k][c] = g.lattice_constant * g.bravais[k][c];
        const double extent = g.bounds_max[k] - g.bounds_min[k];
        if( radius > 0 )
        {
            if( extent > 0 )
                span[k] = std::min( g.n_cells[k] - 1, (int)( 1.1 * radius * g.n_cells[k] / extent ) );
        }
        else
            span[k] = g.n_cells[k] - 1;
        if( std::sqrt( tv[k][0] * tv[k][0] + tv[k][1] * tv[k][1] + tv[k][2] * tv[k][2] ) == 0.0 )
            span[k] = 0;
    }
    const double eps = std::numeric_limits<double>::epsilon();
    for( int ia = 0; ia < g.nb; ++ia )
    {
        double ri[3];
        g.basis_position( ia, ri );
        for( int ta = -span[0]; ta <= span[0]; ++ta )
            for( int tb = -span[1]; tb <= span[1]; ++tb )
                for( int tc = -span[2]; tc <= span[2]; ++tc )
                    for( int ja = 0; ja < g.nb; ++ja )
                    {
                        double rj[3];
                        g.basis_position( ja, rj );
                        for( int c = 0; c < 3; ++c )
                            rj[c] = ( ( rj[c] + ta * tv[0][c] ) + tb * tv[1][c] ) + tc * tv[2][c];
                        const double d0 = ri[0] - rj[0], d1 = ri[1] - rj[1], d2 = ri[2] - rj[2];
                        const double d  = std::sqrt( d0 * d0 + d1 * d1 + d2 * d2 );
                        if( !( d < radius && d > eps ) )
                            continue;
                        pairs.push_back( Pair{ ia, ja, { ta, tb, tc } } );
                        const double r0 = rj[0] - ri[0], r1 = rj[1] - ri[1], r2 = rj[2] - ri[2];
                        const double m  = std::sqrt( r0 * r0 + r1 * r1 + r2 * r2 );
                        dist.push_back( m );
                        normal.push_back( m > 0 ? Vec3{ r0 / m, r1 / m, r2 / m } : Vec3{ 0, 0, 0 } );
                    }
    }
}

void Hamiltonian::update_interactions( const Geometry & g )
{
    exch_pairs.clear();
    exch_J.clear();
    exch_shell.clear();
    if( !exchange_shell_J.empty() )
    {
        neighbours_in_shells( g, (int)exchange_shell_J.size(), true, exch_pairs, exch_shell );
        for( size_t k = 0; k < exch_pairs.size(); ++k )
            exch_J.push_back( exchange_shell_J[exch_shell[k]] );
    }
    else
    {
        exch_pairs = exch_pairs_in;
        exch_J     = exch_J_in;
        for( size_t k = 0; k < exch_pairs_in.size(); ++k )
            exch_shell.push_back( (int)k );
        for( size_t k = 0; k < exch_pairs_in.size(); ++k )
        {
            const Pair & p = exch_pairs_in[k];
            exch_pairs.push_back( Pair{ p.j, p.i, { -p.t[0], -p.t[1], -p.t[2] } } );
            exch_J.push_back( exch_J_in[k] );
            exch_shell.push_back( (int)k );
        }
    }
    dmi_pairs.clear();
    dmi_D.clear();
    dmi_n.clear();
    dmi_shell.clear();
    if( !dmi_shell_D.empty() )
    {
        neighbours_in_shells( g, (int)dmi_shell_D.size(), true, dmi_pairs, dmi_shell );
        for( size_t k = 0; k < dmi_pairs.size(); ++k )
        {
            dmi_n.push_back( dmi_normal_from_pair( g, dmi_pairs[k], dmi_chirality ) );
            dmi_D.push_back( dmi_shell_D[dmi_shell[k]] );
        }
    }
    else
    {
        dmi_pairs = dmi_pairs_in;
        dmi_D     = dmi_D_in;
        dmi_n     = dmi_n_in;
        for( size_t k = 0; k < dmi_pairs_in.size(); ++k )
            dmi_shell.push_back( (int)k );
        for( size_t k = 0; k < dmi_pairs_in.size(); ++k )
        {
            const Pair & p = dmi_pairs_in[k];
            dmi_pairs.push_back( Pair{ p.j, p.i, { -p.t[0], -p.t[1], -p.t[2] } } );
            dmi_D.push_back( dmi_D_in[k] );
            dmi_n.push_back( Vec3{ -dmi_n_in[k][0], -dmi_n_in[k][1], -dmi_n_in[k][2] } );
            dmi_shell.push_back( (int)k );
        }
    }
    ddi_pairs.clear();
    ddi_dist.clear();
    ddi_n.clear();
    if( ddi_method == DDI_CUTOFF && ddi_radius >= 0 )
        pairs_in_radius( g, ddi_radius, ddi_pairs, ddi_dist, ddi_n );
    ddi_bc = bc;
    ++version;
    ++ddi_version;
}

bool Hamiltonian::quadruplets_in_gradient_and_energy() const
{
    return !quads.empty()
           && ( has_zeeman() || !aniso_idx.empty() || !site_K.empty() || !cubic_idx.empty() || !exch_pairs.empty()
                || !dmi_pairs.empty() || ddi_method != DDI_NONE );
}

bool Hamiltonian::has_zeeman() const
{
    return std::fabs( field_magnitude ) > 1e-60;
}

// ---------------------------------------------------------------------------
// Input file
// ---------------------------------------------------------------------------
ConfigFile::ConfigFile( const std::string & path )
{
    std::ifstream f( path, std::ios::in | std::ios::binary );
    if( !f.is_open() )
        return;
    opened = true;
    std::string line;
    while( std::getline( f, line ) )
    {
        line.erase( std::remove_if( line.begin(), line.end(), []( char ch ) { return ch == '|' || ch == '+'; } ),
                    line.end() );
        auto pos = line.find( '#' );
        if( pos == 0 )
            continue;
        if( pos != std::string::npos )
            line.erase( pos );
        if( !line.empty() && line.back() == '\r' )
            line.pop_back();
        lines.push_back( line );
    }
}

bool ConfigFile::find( const std::string & keyword )
{
    std::string kw = keyword;
    std::transform( kw.begin(), kw.end(), kw.begin(), ::tolower );
    for( size_t k = 0; k < lines.size(); ++k )
    {
        std::string l = lines[k];
        std::transform( l.begin(), l.end(), l.begin(), ::tolower );
        if( l.compare( 0, kw.size(), kw ) == 0 )
        {
            iss.clear();
            iss.str( lines[k] );
            std::string dump;
            std::istringstream kws( keyword );
            while( kws >> dump )
                iss >> dump;
            cursor = k + 1;
            return true;
        }
    }
    return false;
}

bool ConfigFile::next_line()
{
    if( cursor >= lines.size() )
        return false;
    iss.clear();
    iss.str( lines[cursor++] );
    return true;
}

long duration_from_string( const std::string & s )
{
    // "h:m:s" -> seconds (utility/Timing.cpp DurationFromString)
    long h = 0, m = 0, sec = 0;
    char c1 = 0, c2 = 0;
    std::istringstream is( s );
    if( is >> h >> c1 >> m >> c2 >> sec )
        return h * 3600 + m * 60 + sec;
    return 0;
}

std::shared_ptr<Geometry> geometry_from_config( const std::string & path )
{
    auto g = std::make_shared<Geometry>();
    if( !path.empty() )
    {
        ConfigFile f( path );
        if( !f.ok() )
            log_message( 2, "Could not open config file \"" + path + "\"; using defaults" );
        else
        {
            f.read( g->lattice_constant, "lattice_constant" );
            if( f.find( "bravais_lattice" ) )
            {
                std::string t;
                f.stream() >> t;
                std::transform( t.begin(), t.end(), t.begin(), ::tolower );
                const double h = 0.5 * std::sqrt( 3.0 );
                if( t == "sc" )
                {
                    g->bravais_type = BravaisType::SC;
                    g->bravais      = { Vec3{ 1, 0, 0 }, Vec3{ 0, 1, 0 }, Vec3{ 0, 0, 1 } };
                }
                else if( t == "fcc" )
                {
                    g->bravais_type = BravaisType::FCC;
                    g->bravais      = { Vec3{ 0.5, 0.0, 0.5 }, Vec3{ 0.5, 0.5, 0.0 }, Vec3{ 0.0, 0.5, 0.5 } };
                }
                else if( t == "bcc" )
                {
                    g->bravais_type = BravaisType::BCC;
                    g->bravais      = { Vec3{ 0.5, 0.5, -0.5 }, Vec3{ -0.5, 0.5, -0.5 }, Vec3{ 0.5, -0.5, -0.5 } };
                }
                else if( t == "hex2d" || t == "hex2d60" )
                {
                    g->bravais_type = BravaisType::Hex2D;
                    g->bravais      = { Vec3{ h, -0.5, 0 }, Vec3{ h, 0.5, 0 }, Vec3{ 0, 0, 1 } };
                }
                else if( t == "hex2d120" )
                {
                    g->bravais_type = BravaisType::Hex2D;
                    g->bravais      = { Vec3{ 0.5, -h, 0 }, Vec3{ 0.5, h, 0 }, Vec3{ 0, 0, 1 } };
                }
                else
                    log_message( 3, "Bravais lattice \"" + t + "\" unknown. Using simple cubic..." );
            }
            else if( f.find( "bravais_vectors" ) )
            {
                g->bravais_type = BravaisType::Irregular;
                for( int k = 0; k < 3; ++k )
                {
                    f.next_line();
                    f.stream() >> g->bravais[k][0] >> g->bravais[k][1] >> g->bravais[k][2];
                }
            }
            else if( f.find( "bravais_matrix" ) )
            {
                g->bravais_type = BravaisType::Irregular;
                for( int k = 0; k < 3; ++k )
                {
                    f.next_line();
                    f.stream() >> g->bravais[0][k] >> g->bravais[1][k] >> g->bravais[2][k];
                }
            }
            f.read3( g->n_cells, "n_basis_cells" );
            if( f.find( "basis" ) )
            {
                f.next_line();
                int n = 0;
                f.stream() >> n;
                if( n > 0 )
                {
                    g->cell_atoms.assign( n, Vec3{ 0, 0, 0 } );
                    g->mu_basis.assign( n, 1.0 );
                    for( int a = 0; a < n; ++a )
                    {
                        f.next_line();
                        f.stream() >> g->cell_atoms[a][0] >> g->cell_atoms[a][1] >> g->cell_atoms[a][2];
                    }
                }
            }
            if( f.find( "mu_s" ) )
            {
                for( size_t a = 0; a < g->cell_atoms.size(); ++a )
                {
                    double v;
                    if( f.stream() >> v )
                        g->mu_basis[a] = v;
                    else
                        g->mu_basis[a] = g->mu_basis[0];
                }
            }
            else
                log_message( 2, "Keyword 'mu_s' not found. Using Default: 1" );
        }
    }
    g->generate();
    return g;
}

static void normalize_vec( Vec3 & v )
{
    double z = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
    if( z > 0 )
    {
        double n = std::sqrt( z );
        for( auto & x : v )
            x /= n;
    }
}

// Quadruplets_from_File (core/src/io/Dataparser.cpp:521-680): after the count (or from the top of
// a separate file), a header line names the columns (i j da_j db_j dc_j k da_k db_k dc_k l da_l
// db_l dc_l Q, any order, case-insensitive), then one quadruplet per line; values of absent
// columns carry over from the previous line, and rows with Q = 0 are dropped.
static void read_quadruplets( ConfigFile & f, long n, Hamiltonian & h )
{
    f.next_line();
    std::vector<std::string> cols;
    std::string w;
    while( f.stream() >> w )
    {
        std::transform( w.begin(), w.end(), w.begin(), ::tolower );
        cols.push_back( w );
    }
    const char * names[14] = { "i",    "j",    "da_j", "db_j", "dc_j", "k",    "da_k",
                               "db_k", "dc_k", "l",    "da_l", "db_l", "dc_l", "q" };
    int iv[13]             = { 0 };
    double Q               = 0;
    for( long r = 0; r < n && f.next_line(); ++r )
    {
        for( auto & c : cols )
        {
            int which = -1;
            for( int k = 0; k < 14; ++k )
                if( c == names[k] )
                    which = k;
            if( which == 13 )
                f.stream() >> Q;
            else if( which >= 0 )
                f.stream() >> iv[which];
            else
            {
                std::string dump;
                f.stream() >> dump;
            }
        }
        if( Q != 0 )
        {
            Quadruplet q{};
            q.i = iv[0];
            q.j = iv[1];
            q.k = iv[5];
            q.l = iv[9];
            for( int c = 0; c < 3; ++c )
            {
                q.dj[c] = iv[2 + c];
                q.dk[c] = iv[6 + c];
                q.dl[c] = iv[10 + c];
            }
            h.quads.push_back( q );
            h.quad_Q.push_back( Q );
        }
    }
}

std::shared_ptr<Hamiltonian> hamiltonian_from_config( const std::string & path, const Geometry & g )
{
    auto h = std::make_shared<Hamiltonian>();
    double B = 0, K = 0, K4 = 0;
    Vec3 Bn{ 0, 0, 1 }, Kn{ 0, 0, 1 };
    std::string ham_type = "heisenberg_neighbours";
    int n_shells_exch = 0, n_shells_dmi = 0;
    std::vector<double> jij, dij;
    std::string ddi = "none";
    if( !path.empty() )
    {
        ConfigFile f( path );
        if( f.ok() )
        {
            f.read( ham_type, "hamiltonian" );
            std::array<int, 3> bci{ 0, 0, 0 };
            f.read3( bci, "boundary_conditions" );
            for( int c = 0; c < 3; ++c )
                h->bc[c] = bci[c] != 0 ? 1 : 0;
            f.read( B, "external_field_magnitude" );
            f.read3( Bn, "external_field_normal" );
            normalize_vec( Bn );
            if( std::sqrt( Bn[0] * Bn[0] + Bn[1] * Bn[1] + Bn[2] * Bn[2] ) < 1e-8 )
                Bn = { 0, 0, 1 };
            f.read( K, "anisotropy_magnitude" );
            f.read3( Kn, "anisotropy_normal" );
            normalize_vec( Kn );
            f.read( K4, "cubic_anisotropy_magnitude" );
            if( ham_type == "heisenberg_pairs" )
            {
                h->from_shells = false;
                if( f.find( "n_interaction_pairs" ) )
                {
                    int n = 0;
                    f.stream() >> n;
                    f.next_line();
                    std::vector<std::string> cols;
                    std::string w;
                    while( f.stream() >> w )
                    {
                        std::transform( w.begin(), w.end(), w.begin(), ::tolower );
                        cols.push_back( w );
                    }
                    for( int r = 0; r < n && f.next_line(); ++r )
                    {
                        int pi = 0, pj = 0, da = 0, db = 0, dc = 0;
                        double J = 0, Dij = 0, D1 = 0, D2 = 0, D3 = 0;
                        bool has_Dij = false;
                        for( auto & cname : cols )
                        {
                            if( cname == "i" )
                                f.stream() >> pi;
                            else if( cname == "j" )
                                f.stream() >> pj;
                            else if( cname == "da" )
                                f.stream() >> da;
                            else if( cname == "db" )
                                f.stream() >> db;
                            else if( cname == "dc" )
                                f.stream() >> dc;
                            else if( cname == "jij" )
                                f.stream() >> J;
                            else if( cname == "dij" )
                            {
                                f.stream() >> Dij;
                                has_Dij = true;
                            }
                            else if( cname == "dijx" )
                                f.stream() >> D1;
                            else if( cname == "dijy" )
                                f.stream() >> D2;
                            else if( cname == "dijz" )
                                f.stream() >> D3;
                            else
                            {
                                std::string dump;
                                f.stream() >> dump;
                            }
                        }
                        // Dataparser.cpp Pairs_from_File: DM vector normalisation
                        double dn = std::sqrt( std::pow( D1, 2 ) + std::pow( D2, 2 ) + std::pow( D3, 2 ) );
                        if( dn != 0 )
                        {
                            D1 = D1 / dn;
                            D2 = D2 / dn;
                            D3 = D3 / dn;
                        }
                        if( !has_Dij )
                            Dij = dn;
                        if( J != 0 )
                        {
                            bool found = false;
                            for( size_t k = 0; k < h->exch_pairs_in.size(); ++k )
                            {
                                auto & p = h->exch_pairs_in[k];
                                if( ( pi == p.i && pj == p.j && da == p.t[0] && db == p.t[1] && dc == p.t[2] )
                                    || ( pi == p.j && pj == p.i && da == -p.t[0] && db == -p.t[1]
                                         && dc == -p.t[2] ) )
                                {
                                    h->exch_J_in[k] += J;
                                    found = true;
                                    break;
                                }
                            }
                            if( !found )
                            {
                                h->exch_pairs_in.push_back( Pair{ pi, pj, { da, db, dc } } );
                                h->exch_J_in.push_back( J );
                            }
                        }
                        if( Dij != 0 )
                        {
                            h->dmi_pairs_in.push_back( Pair{ pi, pj, { da, db, dc } } );
                            h->dmi_D_in.push_back( Dij );
                            h->dmi_n_in.push_back( Vec3{ D1, D2, D3 } );
                        }
                    }
                }
            }
            else
            {
                f.read( n_shells_exch, "n_shells_exchange" );
                jij.assign( n_shells_exch, 0.0 );
                if( n_shells_exch > 0 && f.find( "jij" ) )
                    for( int s = 0; s < n_shells_exch; ++s )
                        f.stream() >> jij[s];
                f.read( n_shells_dmi, "n_shells_dmi" );
                dij.assign( n_shells_dmi, 0.0 );
                if( n_shells_dmi > 0 && f.find( "dij" ) )
                    for( int s = 0; s < n_shells_dmi; ++s )
                        f.stream() >> dij[s];
                f.read( h->dmi_chirality, "dm_chirality" );
            }
            if( f.find( "n_interaction_quadruplets" ) )
            {
                long n = 0;
                f.stream() >> n;
                read_quadruplets( f, n, *h );
            }
            else
            {
                std::string qfile;
                if( f.read( qfile, "interaction_quadruplets_file" ) && !qfile.empty() )
                {
                    ConfigFile qf( qfile );
                    if( !qf.ok() )
                        log_message( 0, "Unable to read interaction quadruplets from \"" + qfile + "\"" );
                    else if( qf.find( "n_interaction_quadruplets" ) )
                    {
                        long n = 0;
                        qf.stream() >> n;
                        read_quadruplets( qf, n, *h );
                    }
                    else
                    {
                        qf.rewind();
                        read_quadruplets( qf, 100000000L, *h );
                    }
                }
            }
            f.read( ddi, "ddi_method" );
            f.read3( h->ddi_images, "ddi_n_periodic_images" );
            int zp = 1;
            if( f.read( zp, "ddi_pb_zero_padding" ) )
                h->ddi_pb_zero_padding = zp != 0;
            f.read( h->ddi_radius, "ddi_radius" );
        }
    }
    h->field_magnitude = B * 0.057883817555;
    h->field_normal    = Bn;
    int nb             = (int)g.cell_atoms.size();
    if( K != 0 )
        for( int a = 0; a < nb; ++a )
        {
            h->aniso_idx.push_back( a );
            h->aniso_K.push_back( K );
            h->aniso_n.push_back( Kn );
        }
    if( K4 != 0 )
        for( int a = 0; a < nb; ++a )
        {
            h->cubic_idx.push_back( a );
            h->cubic_K.push_back( K4 );
        }
    h->exchange_shell_J = jij;
    h->dmi_shell_D      = dij;
    if( ddi == "fft" )
        h->ddi_method = DDI_FFT;
    else if( ddi == "fmm" )
        h->ddi_method = DDI_FMM;
    else if( ddi == "cutoff" )
        h->ddi_method = DDI_CUTOFF;
    else
        h->ddi_method = DDI_NONE;
    h->update_interactions( g );
    return h;
}

LLGParameters llg_parameters_from_config( const std::string & path )
{
    LLGParameters p;
    p.seed = (uint64_t)std::random_device{}();
    if( path.empty() )
        return p;
    ConfigFile f( path );
    if( !f.ok() )
        return p;
    f.read( p.output_file_tag, "output_file_tag" );
    f.read( p.output_folder, "llg_output_folder" );
    f.read( p.output_any, "llg_output_any" );
    f.read( p.output_initial, "llg_output_initial" );
    f.read( p.output_final, "llg_output_final" );
    f.read( p.output_energy_step, "llg_output_energy_step" );
    f.read( p.output_energy_archive, "llg_output_energy_archive" );
    f.read( p.output_energy_spin_resolved, "llg_output_energy_spin_resolved" );
    f.read( p.output_energy_divide_by_nspins, "llg_output_energy_divide_by_nspins" );
    f.read( p.output_energy_add_readability_lines, "llg_output_energy_add_readability_lines" );
    f.read( p.output_configuration_step, "llg_output_configuration_step" );
    f.read( p.output_configuration_archive, "llg_output_configuration_archive" );
    f.read( p.output_configuration_filetype, "llg_output_configuration_filetype" );
    std::string wt;
    if( f.read( wt, "llg_max_walltime" ) )
        p.max_walltime_sec = duration_from_string( wt );
    long seed;
    if( f.read( seed, "llg_seed" ) )
        p.seed = (uint64_t)seed;
    f.read( p.n_iterations, "llg_n_iterations" );
    f.read( p.n_iterations_log, "llg_n_iterations_log" );
    f.read( p.n_iterations_amortize, "llg_n_iterations_amortize" );
    f.read( p.dt, "llg_dt" );
    f.read( p.temperature, "llg_temperature" );
    f.read3( p.temperature_gradient_direction, "llg_temperature_gradient_direction" );
    normalize_vec( p.temperature_gradient_direction );
    f.read( p.temperature_gradient_inclination, "llg_temperature_gradient_inclination" );
    f.read( p.damping, "llg_damping" );
    f.read( p.beta, "llg_beta" );
    f.read( p.stt_use_gradient, "llg_stt_use_gradient" );
    f.read( p.stt_magnitude, "llg_stt_magnitude" );
    f.read3( p.stt_polarisation_normal, "llg_stt_polarisation_normal" );
    normalize_vec( p.stt_polarisation_normal );
    f.read( p.force_convergence, "llg_force_convergence" );
    return p;
}

// Parameters_Method_MC_from_Config, core/src/io/Configparser.cpp:920-980
MCParameters mc_parameters_from_config( const std::string & path )
{
    MCParameters p;
    if( path.empty() )
        return p;
    ConfigFile f( path );
    if( !f.ok() )
        return p;
    f.read( p.output_file_tag, "output_file_tag" );
    f.read( p.output_folder, "mc_output_folder" );
    std::string wt;
    if( f.read( wt, "mc_max_walltime" ) )
        p.max_walltime_sec = duration_from_string( wt );
    long seed;
    if( f.read( seed, "mc_seed" ) )
        p.seed = (uint64_t)seed;
    f.read( p.n_iterations, "mc_n_iterations" );
    f.read( p.n_iterations_log, "mc_n_iterations_log" );
    f.read( p.n_iterations_amortize, "mc_n_iterations_amortize" );
    f.read( p.temperature, "mc_temperature" );
    f.read( p.acceptance_ratio_target, "mc_acceptance_ratio" );
    return p;
}

GNEBParameters gneb_parameters_from_config( const std::string & path )
{
    GNEBParameters p;
    if( path.empty() )
        return p;
    ConfigFile f( path );
    if( !f.ok() )
        return p;
    f.read( p.output_file_tag, "output_file_tag" );
    f.read( p.output_folder, "gneb_output_folder" );
    std::string wt;
    if( f.read( wt, "gneb_max_walltime" ) )
        p.max_walltime_sec = duration_from_string( wt );
    f.read( p.spring_constant, "gneb_spring_constant" );
    f.read( p.force_convergence, "gneb_force_convergence" );
    f.read( p.n_iterations, "gneb_n_iterations" );
    f.read( p.n_iterations_log, "gneb_n_iterations_log" );
    f.read( p.n_iterations_amortize, "gneb_n_iterations_amortize" );
    f.read( p.n_E_interpolations, "gneb_n_energy_interpolations" );
    f.read( p.moving_endpoints, "gneb_moving_endpoints" );
    f.read( p.equilibrium_delta_Rx_left, "gneb_equilibrium_delta_Rx_left" );
    f.read( p.equilibrium_delta_Rx_right, "gneb_equilibrium_delta_Rx_right" );
    f.read( p.translating_endpoints, "gneb_translating_endpoints" );
    // not read by the reference's parser (Configparser.cpp:1031-1057); accepted here so that
    // Method_GNEB's escape_first branch can be switched on
    f.read( p.escape_first, "gneb_escape_first" );
    f.read( p.spring_force_ratio, "gneb_spring_force_ratio" );
    f.read( p.path_shortening_constant, "gneb_path_shortening_constant" );
    return p;
}

} // namespace spirit_hip

namespace spirit_hip
{

// Partner of spin ispin across pair p, or -1 where an open boundary cuts the pair
// (Vectormath.hpp:437-528 with defects off, as the device-side pair_neighbour).
static int pair_partner( const Geometry & g, const std::array<int, 3> & bc, int ispin, const Pair & p )
{
    const int N = g.nb;
    const int n[3] = { g.n_cells[0], g.n_cells[1], g.n_cells[2] };
    if( p.i != ispin % N )
        return -1;
    int cell = ispin / N;
    int idx[3];
    idx[0] = cell % n[0];
    cell /= n[0];
    idx[1] = cell % n[1];
    idx[2] = cell / n[1];
    for( int d = 0; d < 3; ++d )
    {
        if( std::abs( p.t[d] ) > n[d] )
            return -1;
        int j = idx[d] + p.t[d];
        if( !bc[d] && ( j < 0 || j >= n[d] ) )
            return -1;
        idx[d] = ( ( j % n[d] ) + n[d] ) % n[d];
    }
    return p.j + N * ( idx[0] + n[0] * ( idx[1] + n[1] * idx[2] ) );
}

std::vector<HessianEntry> embedding_hessian( const Geometry & g, const Hamiltonian & h )
{
    std::vector<HessianEntry> t;
    const int N = g.nb;
    auto push = [&t]( long i, long j, double v ) { t.push_back( HessianEntry{ i, j, v } ); };

    // single-spin terms; as the reference, every (alpha, beta) entry of the on-site block
    // -2 K n n^T is added at (3i+alpha, 3i+alpha), so the diagonal holds the row sums
    for( int icell = 0; icell < g.n_cells_total; ++icell )
        for( size_t k = 0; k < h.aniso_idx.size(); ++k )
        {
            const long ispin = (long)icell * N + h.aniso_idx[k];
            for( int a = 0; a < 3; ++a )
                for( int b = 0; b < 3; ++b )
                    push( 3 * ispin + a, 3 * ispin + a, -2.0 * h.aniso_K[k] * h.aniso_n[k][a] * h.aniso_n[k][b] );
        }
    // per-site K_i along one axis: treated as the uniaxial term above
    for( size_t ispin = 0; ispin < h.site_K.size(); ++ispin )
        for( int a = 0; a < 3; ++a )
            for( int b = 0; b < 3; ++b )
                push( 3 * (long)ispin + a, 3 * (long)ispin + a, -2.0 * h.site_K[ispin] * h.site_n[a] * h.site_n[b] );

    // exchange and DMI over the redundant pair tables (each pair is listed in both
    // directions, so no transposed entries are added)
    for( int icell = 0; icell < g.n_cells_total; ++icell )
        for( size_t k = 0; k < h.exch_pairs.size(); ++k )
        {
            const int ispin = h.exch_pairs[k].i + icell * N;
            const int jspin = pair_partner( g, h.bc, ispin, h.exch_pairs[k] );
            if( jspin < 0 )
                continue;
            for( int a = 0; a < 3; ++a )
                push( 3L * ispin + a, 3L * jspin + a, -h.exch_J[k] );
        }
    for( int icell = 0; icell < g.n_cells_total; ++icell )
        for( size_t k = 0; k < h.dmi_pairs.size(); ++k )
        {
            const int ispin = h.dmi_pairs[k].i + icell * N;
            const int jspin = pair_partner( g, h.bc, ispin, h.dmi_pairs[k] );
            if( jspin < 0 )
                continue;
            const long i = 3L * ispin, j = 3L * jspin;
            const double D = h.dmi_D[k];
            const Vec3 & n = h.dmi_n[k];
            // d/ds_i of -D n.(s_i x s_j) = -D (s_j x n): the cross-product matrix of n
            push( i + 2, j + 1, D * n[0] );
            push( i + 1, j + 2, -D * n[0] );
            push( i, j + 2, D * n[1] );
            push( i + 2, j, -D * n[1] );
            push( i + 1, j, D * n[2] );
            push( i, j + 1, -D * n[2] );
        }

    // column-major order, duplicates summed in the order they were added
    std::stable_sort( t.begin(), t.end(), []( const HessianEntry & x, const HessianEntry & y ) {
        return x.col != y.col ? x.col < y.col : x.row < y.row;
    } );
    std::vector<HessianEntry> out;
    out.reserve( t.size() );
    for( const auto & e : t )
    {
        if( !out.empty() && out.back().row == e.row && out.back().col == e.col )
            out.back().value += e.value;
        else
            out.push_back( e );
    }
    return out;
}

} // namespace spirit_hip
