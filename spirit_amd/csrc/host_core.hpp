// Host-side data model of the engine: lattice geometry, neighbour tables,
// the Heisenberg Hamiltonian's parameters and the input-file reader.
#pragma once

#include <array>
#include <cstdint>
#include <memory>
#include <mutex>
#include <sstream>
#include <string>
#include <vector>

namespace spirit_hip
{

using Vec3 = std::array<double, 3>;

enum class BravaisType
{
    Irregular = 0,
    Rectilinear,
    SC,
    Hex2D,
    Hex2D60,
    Hex2D120,
    HCP,
    BCC,
    FCC
};

// Lattice, core/src/data/Geometry.cpp:22-161
struct Geometry
{
    std::array<Vec3, 3> bravais{ Vec3{ 1, 0, 0 }, Vec3{ 0, 1, 0 }, Vec3{ 0, 0, 1 } };
    std::array<int, 3> n_cells{ 100, 100, 1 };
    std::vector<Vec3> cell_atoms{ Vec3{ 0, 0, 0 } }; // fractional coordinates
    std::vector<double> mu_basis{ 1.0 };           // mu_s per basis atom
    double lattice_constant = 1.0;
    BravaisType bravais_type = BravaisType::SC;

    // derived
    int nb            = 1;
    int nos           = 0;
    int n_cells_total = 0;
    std::vector<double> positions; // AoS nos*3
    std::vector<double> mu_s;      // per spin
    Vec3 center{ 0, 0, 0 }, bounds_min{ 0, 0, 0 }, bounds_max{ 0, 0, 0 };
    Vec3 cell_bounds_min{ 0, 0, 0 }, cell_bounds_max{ 0, 0, 0 };
    int dimensionality = 2;

    void generate();
    void basis_position( int iatom, double * out ) const;
};

struct Pair
{
    int i, j;
    int t[3];
};

// Four-spin interaction <ij><kl> (Vectormath_Defines.hpp:82-86): basis atoms i, j, k, l and
// the translations of j, k and l from i's cell.
struct Quadruplet
{
    int i, j, k, l;
    int dj[3], dk[3], dl[3];
};

// Neighbours.cpp:15-156, 249-286
std::vector<double> shell_radii( const Geometry & g, int n_shells );
void neighbours_in_shells( const Geometry & g, int n_shells, bool redundant, std::vector<Pair> & pairs,
                           std::vector<int> & shells );
Vec3 dmi_normal_from_pair( const Geometry & g, const Pair & p, int chirality );
// Neighbours::Get_Pairs_in_Radius (Neighbours.cpp:158-246) with DDI_from_Pair (:288-307): every
// pair closer than `radius` (none for |radius| < 1e-6), its distance and unit vector r_ij / |r_ij|
void pairs_in_radius( const Geometry & g, double radius, std::vector<Pair> & pairs, std::vector<double> & dist,
                      std::vector<Vec3> & normal );

enum DdiMethod
{
    DDI_NONE   = 0,
    DDI_FFT    = 1,
    DDI_FMM    = 2,
    DDI_CUTOFF = 3
};

// Heisenberg Hamiltonian parameters, core/include/engine/Hamiltonian_Heisenberg.hpp
struct Hamiltonian
{
    std::array<int, 3> bc{ 0, 0, 0 };
    double field_magnitude = 0; // stored times mu_B, as the reference (Hamiltonian_Heisenberg.cpp:35)
    Vec3 field_normal{ 0, 0, 1 };
    std::vector<int> aniso_idx;
    std::vector<double> aniso_K;
    std::vector<Vec3> aniso_n;
    std::vector<int> cubic_idx;
    std::vector<double> cubic_K;
    std::vector<double> site_K; // per-site uniaxial K_i (extension), along site_n
    Vec3 site_n{ 0, 0, 1 };

    // shell description (heisenberg_neighbours) ...
    bool from_shells = true;
    std::vector<double> exchange_shell_J;
    std::vector<double> dmi_shell_D;
    int dmi_chirality = 1;
    // ... or explicit pairs (heisenberg_pairs)
    std::vector<Pair> exch_pairs_in;
    std::vector<double> exch_J_in;
    std::vector<Pair> dmi_pairs_in;
    std::vector<double> dmi_D_in;
    std::vector<Vec3> dmi_n_in;

    // derived redundant tables, Update_Interactions (Hamiltonian_Heisenberg.cpp:101-198)
    std::vector<Pair> exch_pairs;
    std::vector<double> exch_J;
    std::vector<int> exch_shell; // shell of each pair (input pair index for pair lists)
    std::vector<Pair> dmi_pairs;
    std::vector<double> dmi_D;
    std::vector<Vec3> dmi_n;
    std::vector<int> dmi_shell;

    // quadruplets with their magnitudes Q (n_interaction_quadruplets / interaction_quadruplets_file)
    std::vector<Quadruplet> quads;
    std::vector<double> quad_Q;

    int ddi_method = DDI_NONE;
    std::array<int, 3> ddi_images{ 4, 4, 4 };
    bool ddi_pb_zero_padding = true;
    double ddi_radius        = 0;
    // the cutoff method's pairs (ddi_method == DDI_CUTOFF with ddi_radius >= 0), from
    // update_interactions (Hamiltonian_Heisenberg.cpp:178-190)
    std::vector<Pair> ddi_pairs;
    std::vector<double> ddi_dist;
    std::vector<Vec3> ddi_n;

    // boundary conditions as seen by the last update_interactions (the DDI padding is
    // prepared there, Hamiltonian_Heisenberg.cpp:1501-1520, so later BC changes keep it)
    std::array<int, 3> ddi_bc{ 0, 0, 0 };

    // bumped whenever something changes that invalidates device-side copies
    uint64_t version     = 0;
    uint64_t ddi_version = 0;

    void update_interactions( const Geometry & g );
    bool has_zeeman() const;
    // Gradient_and_Energy adds the quadruplets only when idx_quadruplet > 0
    // (Hamiltonian_Heisenberg.cpp:754), i.e. when another contribution precedes them in the
    // list of Update_Energy_Contributions (:200-260)
    bool quadruplets_in_gradient_and_energy() const;
};

struct LLGParameters
{
    double dt                = 1e-3;
    double damping           = 0.3;
    double beta              = 0;
    double temperature       = 0;
    Vec3 temperature_gradient_direction{ 1, 0, 0 };
    double temperature_gradient_inclination = 0;
    bool stt_use_gradient    = true;
    double stt_magnitude     = 0;
    Vec3 stt_polarisation_normal{ 1, 0, 0 };
    bool direct_minimization = false;
    long n_iterations        = 1000000;
    long n_iterations_log    = 1000;
    long n_iterations_amortize = 1;
    long max_walltime_sec    = 0;
    double force_convergence = 1e-10;
    uint64_t seed            = 2006;
    std::string output_folder = "output";
    std::string output_file_tag = "<time>";
    bool output_any = false, output_initial = false, output_final = false;
    bool output_energy_step = false, output_energy_archive = false, output_energy_spin_resolved = false;
    bool output_energy_divide_by_nspins = true, output_energy_add_readability_lines = false;
    bool output_configuration_step = false, output_configuration_archive = false;
    int output_configuration_filetype = 3;
};

// Monte Carlo parameters, core/include/data/Parameters_Method_MC.hpp
struct MCParameters
{
    long n_iterations             = 1000000;
    long n_iterations_log         = 1000;
    long n_iterations_amortize    = 1;
    long max_walltime_sec         = 0;
    double temperature            = 0;
    uint64_t seed                 = 2006;
    bool metropolis_random_sample = true;
    bool metropolis_step_cone     = true;
    bool metropolis_cone_adaptive = true;
    double metropolis_cone_angle  = 30; // degrees
    double acceptance_ratio_target = 0.5;
    std::string output_folder = "output";
    std::string output_file_tag = "<time>";
};

struct GNEBParameters
{
    double spring_constant          = 1;
    double spring_force_ratio       = 0;
    double path_shortening_constant = 0;
    int n_E_interpolations          = 10;
    bool moving_endpoints           = false;
    bool translating_endpoints      = false;
    double equilibrium_delta_Rx_left  = 1.0;
    double equilibrium_delta_Rx_right = 1.0;
    bool escape_first               = false;
    long n_iterations               = 1000000;
    long n_iterations_log           = 1000;
    long n_iterations_amortize      = 1;
    long max_walltime_sec           = 0;
    double force_convergence        = 1e-10;
    std::string output_folder = "output";
    std::string output_file_tag = "<time>";
    bool output_any = false, output_initial = false, output_final = false;
    bool output_energies_step = false, output_energies_divide_by_nspins = true;
    bool output_energies_add_readability_lines = false, output_energies_interpolated = false;
    bool output_chain_step = false;
    int output_chain_filetype = 3;
};

// ---------------------------------------------------------------------------
// Input file reader with the reference's keyword semantics
// (core/src/io/Filter_File_Handle.cpp): '#' comments, '|' and '+' removed,
// a keyword matches a line that starts with it (case-insensitive).
// ---------------------------------------------------------------------------
class ConfigFile
{
public:
    explicit ConfigFile( const std::string & path );
    bool ok() const
    {
        return opened;
    }
    // positions the reader after the keyword on its line; returns false if absent
    bool find( const std::string & keyword );
    // next non-comment line
    bool next_line();
    // the next next_line() reads the first line again
    void rewind()
    {
        cursor = 0;
    }
    std::istringstream & stream()
    {
        return iss;
    }
    template<typename T>
    bool read( T & v, const std::string & keyword )
    {
        if( !find( keyword ) )
            return false;
        T tmp;
        if( iss >> tmp )
        {
            v = tmp;
            return true;
        }
        return false;
    }
    template<typename T>
    bool read3( T & v, const std::string & keyword )
    {
        if( !find( keyword ) )
            return false;
        T tmp = v;
        if( iss >> tmp[0] >> tmp[1] >> tmp[2] )
        {
            v = tmp;
            return true;
        }
        return false;
    }

private:
    bool opened = false;
    std::vector<std::string> lines;
    size_t cursor = 0;
    std::istringstream iss;
};

// Builders from an input file, restating core/src/io/Configparser.cpp
// (Geometry_from_Config :271-567, Parameters_Method_LLG_from_Config :701-824,
//  Parameters_Method_GNEB_from_Config, Hamiltonian_Heisenberg_from_Config :1256-1630).
std::shared_ptr<Geometry> geometry_from_config( const std::string & path );
std::shared_ptr<Hamiltonian> hamiltonian_from_config( const std::string & path, const Geometry & g );
LLGParameters llg_parameters_from_config( const std::string & path );
MCParameters mc_parameters_from_config( const std::string & path );
GNEBParameters gneb_parameters_from_config( const std::string & path );
long duration_from_string( const std::string & s );

// The 3N x 3N embedding Hessian of the quadratic terms (Hamiltonian_Heisenberg::Sparse_Hessian,
// core/src/engine/Hamiltonian_Heisenberg.cpp:1277-1371): uniaxial anisotropy, exchange and DMI;
// Zeeman, cubic anisotropy and DDI contribute nothing. Entries are (row, col, value) with
// duplicates summed in insertion order and sorted column-major, rows ascending, the order in
// which a column-major sparse matrix built from the same triplets iterates its entries.
// Entries whose value is zero stay in the list (they are part of the sparsity pattern).
struct HessianEntry
{
    long row, col;
    double value;
};
std::vector<HessianEntry> embedding_hessian( const Geometry & g, const Hamiltonian & h );

// logging
void log_message( int level, const std::string & msg );
extern int g_log_console_level;
extern bool g_quiet;
extern int g_n_errors;
extern int g_n_warnings;

} // namespace spirit_hip
