// IO_* entry points: spin configurations and chains in OVF 2.0 files.
//
// Restated reference code (paths relative to the reference checkout):
//   core/src/Spirit/IO.cpp:90-812 (IO_Positions_Write, IO_N_Images_In_File, IO_Image_Read /
//   _Write / _Append, IO_Chain_Read / _Write / _Append), core/src/io/OVF_File.cpp:12-39
#include "../../include/Spirit/IO.h"
#include "api_common.hpp"
#include "ovf.hpp"

#include <algorithm>
#include <charconv>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <ctime>

using namespace spirit_hip;

namespace
{

// IO::OVF_Segment( system ), OVF_File.cpp:12-39, with the title and value labels IO.cpp sets
ovf::Segment image_segment( const Image & img, const char * comment, const char * labels )
{
    const Geometry & g = *img.geometry;
    ovf::Segment s;
    s.title       = std::string( "SPIRIT Version " ) + Spirit_Version_Full();
    s.comment     = comment ? comment : "";
    s.valuedim    = 3;
    s.valuelabels = labels;
    s.valueunits  = "none none none";
    s.meshtype    = "rectangular";
    s.meshunit    = "nm";
    s.n_cells[0]  = g.n_cells[0] * g.nb;
    s.n_cells[1]  = g.n_cells[1];
    s.n_cells[2]  = g.n_cells[2];
    for( int k = 0; k < 3; ++k )
    {
        s.bounds_min[k] = g.bounds_min[k] * 0.1;
        s.bounds_max[k] = g.bounds_max[k] * 0.1;
        s.origin[k]     = 0;
        s.step_size[k]  = g.lattice_constant * g.bravais[k][k] * 0.1;
    }
    return s;
}

// The format written for a requested one.  Spins and positions are doubles, and the
// reference's double-precision writer turns every binary request into binary in the data's
// own precision (OVF_File.cpp:127-135, ovf.cpp:371-373): Binary 8, also for bin4.
int written_format( int format )
{
    if( format < IO_Fileformat_OVF_bin || format > IO_Fileformat_OVF_csv )
        throw api_error( 2, "Invalid file format index " + std::to_string( format ) );
    if( format == IO_Fileformat_OVF_bin4 || format == IO_Fileformat_OVF_bin8 )
        return IO_Fileformat_OVF_bin;
    return format;
}

void warn_extension( const char * file, int idx_image, int idx_chain )
{
    const std::string f = file;
    if( f.size() < 4 || f.substr( f.size() - 4 ) != ".ovf" )
        log_message( 3, "The file \"" + f
                            + "\" is written in OVF format but has different extension. It is recommend to use the "
                              "appropriate \".ovf\" extension" );
    (void)idx_image;
    (void)idx_chain;
}

// segment data into an image's spins: the first min(N, nos) rows, each normalised, vacancies
// (|s| < 1e-5) set to +z (IO.cpp:231-283)
void read_into( Image & img, const std::string & file, int idx_infile, int n_segments )
{
    ovf::Segment seg;
    std::vector<double> data;
    ovf::read_segment( file, idx_infile, seg, data );
    long N = seg.rows();
    if( N != img.nos() )
        log_message( 3, "OVF file \"" + file + "\": segment " + std::to_string( idx_infile + 1 ) + "/"
                            + std::to_string( n_segments ) + " contains " + std::to_string( N )
                            + " spins while the system contains " + std::to_string( img.nos() ) + "." );
    if( seg.valuedim != 3 )
        throw api_error( 2, "Segment " + std::to_string( idx_infile + 1 ) + "/" + std::to_string( n_segments )
                                + " in OVF file \"" + file + "\" should have 3 columns, but only has "
                                + std::to_string( seg.valuedim ) + ". Will not read." );
    N = std::min<long>( N, img.nos() );
    for( long i = 0; i < N; ++i )
    {
        double v[3] = { data[3 * i], data[3 * i + 1], data[3 * i + 2] };
        double n    = std::sqrt( v[0] * v[0] + v[1] * v[1] + v[2] * v[2] );
        if( n < 1e-5 )
            v[0] = 0, v[1] = 0, v[2] = 1;
        else
            v[0] /= n, v[1] /= n, v[2] /= n;
        for( int c = 0; c < 3; ++c )
            img.spins[3 * i + c] = v[c];
    }
}

} // namespace

extern "C" void IO_Positions_Write( State * state, const char * file, int format, const char * comment,
                                    int idx_image, int idx_chain ) noexcept
try
{
    std::shared_ptr<Image> image;
    std::shared_ptr<Chain> chain;
    from_indices( state, idx_image, idx_chain, image, chain );
    format = written_format( format );
    warn_extension( file, idx_image, idx_chain );
    std::lock_guard<std::recursive_mutex> lk( image->mtx );
    auto seg = image_segment( *image, comment, "position_x position_y position_z" );
    ovf::write_segment( file, seg, image->geometry->positions.data(), format, false );
}
catch( ... )
{
    handle_exception( idx_image, idx_chain );
}

extern "C" int IO_N_Images_In_File( State * state, const char * file, int idx_image, int idx_chain ) noexcept
try
{
    (void)state;
    auto info = ovf::inspect( file );
    if( info.is_ovf )
        return info.n_segments;
    log_message( 3, std::string( "File \"" ) + file + "\" is not OVF. Cannot measure number of images." );
    return -1;
}
catch( ... )
{
    handle_exception( idx_image, idx_chain );
    return -1;
}

extern "C" void IO_Image_Read( State * state, const char * file, int idx_image_infile, int idx_image_inchain,
                               int idx_chain ) noexcept
try
{
    std::shared_ptr<Image> image;
    std::shared_ptr<Chain> chain;
    from_indices( state, idx_image_inchain, idx_chain, image, chain );
    auto info = ovf::inspect( file );
    if( !info.is_ovf )
        throw api_error( 2, std::string( "File \"" ) + file + "\" does not seem to be in valid OVF format. Message: "
                                + info.message + "." );
    std::lock_guard<std::recursive_mutex> lk( image->mtx );
    sync_image_host( state, idx_image_inchain );
    read_into( *image, file, idx_image_infile, info.n_segments );
    notify_image_changed( state, idx_image_inchain, true, false, false );
    log_message( 5, std::string( "Read image from file \"" ) + file + "\"" );
}
catch( ... )
{
    handle_exception( idx_image_inchain, idx_chain );
}

static void image_write( State * state, const char * file, int format, const char * comment, int idx_image,
                         int idx_chain, bool append )
{
    std::shared_ptr<Image> image;
    std::shared_ptr<Chain> chain;
    from_indices( state, idx_image, idx_chain, image, chain );
    format = written_format( format );
    if( !append )
        warn_extension( file, idx_image, idx_chain );
    else
    {
        auto info = ovf::inspect( file );
        if( info.found && !info.is_ovf )
            throw api_error( 2, std::string( "Cannot append to non-OVF file \"" ) + file + "\"" );
    }
    std::lock_guard<std::recursive_mutex> lk( image->mtx );
    sync_image_host( state, idx_image );
    auto seg = image_segment( *image, comment, "spin_x spin_y spin_z" );
    ovf::write_segment( file, seg, image->spins.data(), format, append );
}

extern "C" void IO_Image_Write( State * state, const char * file, int format, const char * comment, int idx_image,
                                int idx_chain ) noexcept
try
{
    image_write( state, file, format, comment, idx_image, idx_chain, false );
}
catch( ... )
{
    handle_exception( idx_image, idx_chain );
}

extern "C" void IO_Image_Append( State * state, const char * file, int format, const char * comment, int idx_image,
                                 int idx_chain ) noexcept
try
{
    image_write( state, file, format, comment, idx_image, idx_chain, true );
}
catch( ... )
{
    handle_exception( idx_image, idx_chain );
}

// IO.cpp:439-649.  The images from start_image_infile .. end_image_infile go to the chain
// images insert_idx, insert_idx + 1, ...; the chain grows (copies of its last image) when
// they do not fit.
extern "C" void IO_Chain_Read( State * state, const char * file, int start_image_infile, int end_image_infile,
                               int insert_idx, int idx_chain ) noexcept
try
{
    int idx0 = insert_idx;
    std::shared_ptr<Image> image;
    std::shared_ptr<Chain> chain;
    from_indices( state, idx0, idx_chain, image, chain );
    if( insert_idx < 0 )
        insert_idx = 0;
    const int noi = chain->noi();
    if( insert_idx > noi )
        throw api_error( 2, "IO_Chain_Read: Tried to start reading chain on invalid index (insert_idx="
                                + std::to_string( insert_idx ) + ", but chain has " + std::to_string( noi )
                                + " images)" );
    auto info = ovf::inspect( file );
    if( !info.is_ovf )
        throw api_error( 2, std::string( "IO_Chain_Read: File \"" ) + file + "\" is not in OVF format: "
                                + info.message );
    const int noi_infile = info.n_segments;
    if( start_image_infile < 0 )
        start_image_infile = 0;
    if( end_image_infile < 0 )
        end_image_infile = noi_infile - 1;
    if( end_image_infile < start_image_infile || end_image_infile >= noi_infile )
    {
        log_message( 3, "IO_Chain_Read: specified invalid reading range (start_image_infile="
                            + std::to_string( start_image_infile ) + ", end_image_infile="
                            + std::to_string( end_image_infile ) + "). Set to read entire file \"" + file + "\" ("
                            + std::to_string( noi_infile ) + " images)." );
        end_image_infile = noi_infile - 1;
    }
    if( start_image_infile >= noi_infile )
        throw api_error( 2, "Specified starting index " + std::to_string( start_image_infile ) + ", but file \""
                                + file + "\" contains only " + std::to_string( noi_infile ) + " images." );
    const int noi_to_read = end_image_infile - start_image_infile + 1;
    const int noi_to_add  = noi_to_read - ( noi - insert_idx );
    if( noi_to_add > 0 )
    {
        Chain_Image_to_Clipboard( state, noi - 1, idx_chain );
        Chain_Set_Length( state, noi + noi_to_add, idx_chain );
    }
    sync_all_hosts( state );
    for( int k = 0; k < noi_to_read; ++k )
    {
        const int i = insert_idx + k;
        auto & img  = *chain->images[i];
        std::lock_guard<std::recursive_mutex> lk( img.mtx );
        read_into( img, file, start_image_infile + k, noi_infile );
        notify_image_changed( state, i, true, false, false );
    }
    State_Update( state );
    Chain_Setup_Data( state, idx_chain );
    log_message( 5, std::string( "IO_Chain_Read: Read chain from file \"" ) + file + "\"" );
}
catch( ... )
{
    handle_exception( insert_idx, idx_chain );
}

// IO.cpp:651-812: one segment per image, "Image i of noi. comment" (i from 1)
static void chain_write( State * state, const char * file, int format, const char * comment, int idx_chain,
                         bool append )
{
    int idx_image = 0;
    std::shared_ptr<Image> image;
    std::shared_ptr<Chain> chain;
    from_indices( state, idx_image, idx_chain, image, chain );
    format = written_format( format );
    if( append )
    {
        auto info = ovf::inspect( file );
        if( info.found && !info.is_ovf )
            throw api_error( 2, std::string( "Cannot append to non-OVF file \"" ) + file + "\"" );
    }
    std::lock_guard<std::recursive_mutex> lk( chain->mtx );
    sync_all_hosts( state );
    const int noi = chain->noi();
    for( int i = 0; i < noi; ++i )
    {
        const std::string c = "Image " + std::to_string( i + 1 ) + " of " + std::to_string( noi ) + ". "
                              + ( comment ? comment : "" );
        auto seg = image_segment( *chain->images[i], c.c_str(), "spin_x spin_y spin_z" );
        ovf::write_segment( file, seg, chain->images[i]->spins.data(), format, append || i > 0 );
    }
}

extern "C" void IO_Chain_Write( State * state, const char * file, int format, const char * comment,
                                int idx_chain ) noexcept
try
{
    chain_write( state, file, format, comment, idx_chain, false );
}
catch( ... )
{
    handle_exception( 0, idx_chain );
}

extern "C" void IO_Chain_Append( State * state, const char * file, int format, const char * comment,
                                 int idx_chain ) noexcept
try
{
    chain_write( state, file, format, comment, idx_chain, true );
}
catch( ... )
{
    handle_exception( 0, idx_chain );
}

// Spin configuration files of an LLG run (Method_LLG::Save_Current, Method_LLG.cpp:310-495):
// <folder>/<tag>_Image-<ii>_Spins<suffix>.ovf with the suffixes "-initial", "-final",
// "_<iteration>" (one file per log step) and "-archive" (one segment appended per log step).
// The tag is the run's start time ("<time>"), the given tag, or nothing.  Energy files are
// not written.
void spirit_hip::llg_save_configuration( const Image & img, int idx_image, const std::string & starttime,
                                         long iteration, long n_iterations, double max_torque,
                                         const std::string & solver_full_name, bool initial, bool final )
{
    const LLGParameters & P = *img.llg;
    if( !P.output_any )
        return;
    char s_img[32];
    std::snprintf( s_img, sizeof( s_img ), "%02d", idx_image );
    // iteration zero-padded to floor(log10(n_iterations)) digits
    const int base = n_iterations > 0 ? (int)std::floor( std::log10( (double)n_iterations ) ) : 0;
    char s_iter[64];
    std::snprintf( s_iter, sizeof( s_iter ), "%0*ld", base, iteration );
    std::string tag = P.output_file_tag == "<time>" ? starttime + "_"
                      : P.output_file_tag.empty()   ? std::string()
                                                    : P.output_file_tag + "_";
    const std::string pre = P.output_folder + "/" + tag + "Image-" + s_img + "_Spins";
    char torque[64];
    {
        auto r = std::to_chars( torque, torque + sizeof( torque ) - 1, max_torque );
        *r.ptr = 0;
    }
    const std::string comment = "LLG simulation (" + solver_full_name + " solver)\n# Desc:      Iteration: "
                                + std::to_string( iteration ) + "\n# Desc:      Maximum torque: " + torque;
    auto write = [&]( const std::string & suffix, bool append ) {
        try
        {
            auto seg = image_segment( img, comment.c_str(), "spin_x spin_y spin_z" );
            ovf::write_segment( pre + suffix + ".ovf", seg, img.spins.data(),
                                written_format( P.output_configuration_filetype ), append );
        }
        catch( const std::exception & e )
        {
            log_message( 2, std::string( "LLG output failed: " ) + e.what() );
        }
    };
    if( initial && P.output_initial )
        write( "-initial", false );
    else if( final && P.output_final )
        write( "-final", false );
    if( P.output_configuration_step )
        write( std::string( "_" ) + s_iter, false );
    if( P.output_configuration_archive )
        write( "-archive", true );
}

namespace
{

// fmt's "{:^w}": centred, the odd padding character on the right
std::string centred( const std::string & s, int w )
{
    const int pad = w - (int)s.size();
    if( pad <= 0 )
        return s;
    return std::string( pad / 2, ' ' ) + s + std::string( pad - pad / 2, ' ' );
}

std::string fixed10( double v )
{
    char b[64];
    std::snprintf( b, sizeof( b ), "%.10f", v );
    return b;
}

void put_text( const std::string & text, const std::string & path, bool append )
{
    std::FILE * f = std::fopen( path.c_str(), append ? "ab" : "wb" );
    if( !f )
    {
        log_message( 2, "Unable to write to file \"" + path + "\"" );
        return;
    }
    std::fwrite( text.data(), 1, text.size(), f );
    std::fclose( f );
}

// IO::Write_Energy_Header (Datawriter.cpp:116-160) with the columns "iteration", "E_tot"
std::string energy_header( const Image & img, bool readability )
{
    std::string separator, line;
    for( const char * column : { "iteration", "E_tot" } )
    {
        if( readability )
            separator += "----------------------++";
        line += " " + centred( column, 20 ) + " ||";
    }
    bool first = true;
    for( const auto & pair : img.E_array )
    {
        if( first )
            first = false;
        else
        {
            line += "|";
            if( readability )
                separator += "+";
        }
        line += " " + centred( pair.first, 20 ) + " ";
        if( readability )
            separator += "----------------------";
    }
    line += "\n";
    separator += "\n";
    std::string header = readability ? separator + line + separator : line;
    if( !readability )
        std::replace( header.begin(), header.end(), '|', ' ' );
    return header;
}

// IO::Append_Image_Energy (Datawriter.cpp:162-185)
std::string energy_line( const Image & img, long iteration, bool normalize, bool readability )
{
    const double norm = normalize ? 1.0 / (double)img.nos() : 1.0;
    std::string line  = " " + centred( std::to_string( iteration ), 20 ) + " || "
                       + centred( fixed10( img.E * norm ), 20 ) + " |";
    for( const auto & pair : img.E_array )
        line += "| " + centred( fixed10( pair.second * norm ), 20 ) + " ";
    line += "\n";
    if( !readability )
        std::replace( line.begin(), line.end(), '|', ' ' );
    return line;
}

} // namespace

// The energy files of Method_LLG::Save_Current (Method_LLG.cpp:383-495): "<pre>_Energy<suffix>.txt"
// holds the energy header and one line (image energy and the contributions of its last
// update); a fresh file (initial, final, step) also writes the per-spin energies
// "<pre>_Energy-perSpin<suffix>.ovf" when spin-resolved output is asked for (energies from
// the GPU, contributions updated); archive files get a header once and a line per save.
void spirit_hip::llg_save_energy( Image & img, int idx_image, const std::string & starttime, long iteration,
                                  long n_iterations, bool initial, bool final )
{
    const LLGParameters & P = *img.llg;
    if( !P.output_any )
        return;
    char s_img[32];
    std::snprintf( s_img, sizeof( s_img ), "%02d", idx_image );
    const int base = n_iterations > 0 ? (int)std::floor( std::log10( (double)n_iterations ) ) : 0;
    char s_iter[64];
    std::snprintf( s_iter, sizeof( s_iter ), "%0*ld", base, iteration );
    const std::string tag = P.output_file_tag == "<time>" ? starttime + "_"
                            : P.output_file_tag.empty()   ? std::string()
                                                          : P.output_file_tag + "_";
    const std::string pre = P.output_folder + "/" + tag + "Image-" + s_img + "_Energy";
    const bool normalize  = P.output_energy_divide_by_nspins;
    const bool readable   = P.output_energy_add_readability_lines;
    auto write = [&]( const std::string & suffix, bool append ) {
        const std::string file = pre + suffix + ".txt";
        if( append )
        {
            std::FILE * f = std::fopen( file.c_str(), "rb" );
            if( f )
                std::fclose( f );
            else
                put_text( energy_header( img, readable ), file, false );
            put_text( energy_line( img, iteration, normalize, readable ), file, true );
            return;
        }
        put_text( energy_header( img, readable ), file, false );
        put_text( energy_line( img, iteration, normalize, readable ), file, true );
        if( !P.output_energy_spin_resolved )
            return;
        try
        {
            std::vector<double> terms;
            img.E           = update_image_energy( img, false, &terms );
            const auto list = energy_contribution_list( img );
            const size_t nos = (size_t)img.nos(), dim = 1 + list.size();
            std::vector<double> data( dim * nos, 0.0 );
            for( size_t i = 0; i < nos; ++i )
            {
                double sum = 0;
                for( size_t k = 0; k < list.size(); ++k )
                {
                    const double e       = terms[(size_t)list[k].second * nos + i];
                    data[i * dim + 1 + k] = e;
                    sum += e;
                }
                data[i * dim] = sum;
            }
            ovf::Segment seg = image_segment( img, "", "" );
            char total[64];
            {
                auto r = std::to_chars( total, total + sizeof( total ) - 1, img.E );
                *r.ptr = 0;
            }
            std::string comment = std::string( "Energy per spin. Total=" ) + total + "meV";
            std::string labels  = "Total";
            for( const auto & c : img.E_array )
            {
                char v[64];
                auto r = std::to_chars( v, v + sizeof( v ) - 1, c.second );
                *r.ptr = 0;
                comment += ", " + c.first + "=" + v + "meV";
                labels += " " + c.first;
            }
            seg.comment     = comment;
            seg.valuedim    = (int)dim;
            seg.valuelabels = labels;
            seg.valueunits  = "";
            ovf::write_segment( pre + "-perSpin" + suffix + ".ovf", seg, data.data(),
                                written_format( P.output_configuration_filetype ), false );
        }
        catch( const std::exception & e )
        {
            log_message( 2, std::string( "LLG energy output failed: " ) + e.what() );
        }
    };
    if( initial && P.output_initial )
        write( "-initial", false );
    else if( final && P.output_final )
        write( "-final", false );
    if( P.output_energy_step )
        write( std::string( "_" ) + s_iter, false );
    if( P.output_energy_archive )
        write( "-archive", true );
}

std::string spirit_hip::current_date_time()
{
    const std::time_t t = std::time( nullptr );
    const std::tm tm_s  = *std::localtime( &t );
    char b[64];
    std::strftime( b, sizeof( b ), "%Y-%m-%d_%H-%M-%S", &tm_s );
    return b;
}

