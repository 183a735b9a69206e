// OVF 2.0 files: writer, inspector and reader (host code).  See ovf.hpp.
#include "ovf.hpp"

#include <algorithm>
#include <cctype>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <stdexcept>

namespace spirit_hip
{
namespace ovf
{

// check values of the binary data blocks (float32 1234567.0, float64 123456789012345.0)
static const uint32_t CHECK_4B = 0x4996B438u;
static const uint64_t CHECK_8B = 0x42DC12218377DE40ull;
static const int COUNT_DIGITS  = 6; // zero-padded segment count, rewritten in place on append

long Segment::rows() const
{
    if( meshtype == "irregular" )
        return pointcount;
    return (long)n_cells[0] * n_cells[1] * n_cells[2];
}

// a header number as the reference's writer prints it: the segment holds it as a float,
// formatted as printf %g
static std::string num( double v )
{
    char b[64];
    std::snprintf( b, sizeof( b ), "%g", (double)(float)v );
    return b;
}

static std::string lower( std::string s )
{
    for( auto & c : s )
        c = (char)std::tolower( (unsigned char)c );
    return s;
}

static std::string trim( const std::string & s )
{
    size_t a = s.find_first_not_of( " \t\r\n" ), b = s.find_last_not_of( " \t\r\n" );
    return a == std::string::npos ? std::string() : s.substr( a, b - a + 1 );
}

static std::string segment_text( const Segment & seg, const double * data, int format )
{
    const std::string e = "#\n";
    std::string o;
    o += e;
    o += "# Begin: Segment\n";
    o += "# Begin: Header\n";
    o += e;
    o += "# Title: " + seg.title + "\n";
    o += e;
    o += "# Desc: " + seg.comment + "\n";
    o += e;
    o += "# valuedim: " + std::to_string( seg.valuedim ) + "   ## field dimensionality\n";
    auto unspecified = [&]() {
        std::string u;
        for( int i = 0; i < seg.valuedim; ++i )
            u += " unspecified";
        return u;
    };
    o += seg.valueunits.empty() ? "# valueunits: " + unspecified() + "\n" : "# valueunits: " + seg.valueunits + "\n";
    o += seg.valuelabels.empty() ? "# valuelabels: " + unspecified() + "\n"
                                 : "# valuelabels: " + seg.valuelabels + "\n";
    o += e;
    o += "## Fundamental mesh measurement unit. Treated as a label:\n";
    o += "# meshunit: " + ( seg.meshunit.empty() ? std::string( "unspecified" ) : seg.meshunit ) + "\n";
    o += e;
    const char * ax[3] = { "x", "y", "z" };
    for( int k = 0; k < 3; ++k )
        o += std::string( "# " ) + ax[k] + "min: " + num( seg.bounds_min[k] ) + "\n";
    for( int k = 0; k < 3; ++k )
        o += std::string( "# " ) + ax[k] + "max: " + num( seg.bounds_max[k] ) + "\n";
    o += e;
    const std::string meshtype = seg.meshtype.empty() ? std::string( "rectangular" ) : seg.meshtype;
    o += "# meshtype: " + meshtype + "\n";
    if( meshtype == "rectangular" )
    {
        for( int k = 0; k < 3; ++k )
            o += std::string( "# " ) + ax[k] + "base: " + num( seg.origin[k] ) + "\n";
        for( int k = 0; k < 3; ++k )
            o += std::string( "# " ) + ax[k] + "stepsize: " + num( seg.step_size[k] ) + "\n";
        for( int k = 0; k < 3; ++k )
            o += std::string( "# " ) + ax[k] + "nodes: " + std::to_string( seg.n_cells[k] ) + "\n";
    }
    else if( meshtype == "irregular" )
        o += "# pointcount: " + std::to_string( seg.pointcount ) + "\n";
    else
        throw std::runtime_error( "OVF: invalid meshtype \"" + meshtype + "\"" );
    Segment s2 = seg;
    s2.meshtype = meshtype;
    const long n_rows = s2.rows();
    const int n_cols  = seg.valuedim;
    if( (long)n_cols * n_rows <= 0 )
        throw std::runtime_error( "OVF: segment without data (valuedim * rows <= 0)" );
    o += e;
    o += "# End: Header\n";
    o += e;
    if( format == FORMAT_BIN )
        format = FORMAT_BIN8;
    std::string type;
    if( format == FORMAT_BIN8 )
        type = "Binary 8";
    else if( format == FORMAT_BIN4 )
        type = "Binary 4";
    else if( format == FORMAT_TEXT )
        type = "Text";
    else if( format == FORMAT_CSV )
        type = "CSV";
    else
        throw std::runtime_error( "OVF: invalid format " + std::to_string( format ) );
    o += "# Begin: Data " + type + "\n";
    const long n = n_rows * n_cols;
    if( format == FORMAT_BIN8 )
    {
        // little-endian IEEE doubles (the host's own byte order on x86-64)
        o.append( reinterpret_cast<const char *>( &CHECK_8B ), 8 );
        o.append( reinterpret_cast<const char *>( data ), sizeof( double ) * (size_t)n );
        o += "\n";
    }
    else if( format == FORMAT_BIN4 )
    {
        o.append( reinterpret_cast<const char *>( &CHECK_4B ), 4 );
        std::vector<float> f( (size_t)n );
        for( long i = 0; i < n; ++i )
            f[i] = (float)data[i];
        o.append( reinterpret_cast<const char *>( f.data() ), sizeof( float ) * (size_t)n );
        o += "\n";
    }
    else
    {
        const char * delim = format == FORMAT_CSV ? "," : "";
        char b[64];
        for( long r = 0; r < n_rows; ++r )
        {
            for( int c = 0; c < n_cols; ++c )
            {
                std::snprintf( b, sizeof( b ), "%22.12f%s", data[r * n_cols + c], delim );
                o += b;
            }
            o += "\n";
        }
    }
    o += "# End: Data " + type + "\n";
    o += "# End: Segment\n";
    return o;
}

static std::string read_all( const std::string & file, bool & found )
{
    std::ifstream in( file, std::ios::binary );
    found = (bool)in;
    if( !in )
        return std::string();
    std::ostringstream ss;
    ss << in.rdbuf();
    return ss.str();
}

// offset of the digits of "# Segment count:" in the file header, -1 if absent
static long count_pos( const std::string & text )
{
    const std::string key = "# Segment count:";
    size_t p              = text.find( key );
    if( p == std::string::npos || p > 256 )
        return -1;
    p += key.size();
    while( p < text.size() && text[p] == ' ' )
        ++p;
    return (long)p;
}

FileInfo inspect( const std::string & file )
{
    FileInfo f;
    std::ifstream in( file, std::ios::binary );
    if( !in )
    {
        f.message = "file not found";
        return f;
    }
    f.found = true;
    std::string head( 512, '\0' );
    in.read( &head[0], (std::streamsize)head.size() );
    head.resize( (size_t)in.gcount() );
    size_t eol = head.find( '\n' );
    const std::string first = trim( head.substr( 0, eol ) );
    if( lower( first ) != "# oommf ovf 2.0" )
    {
        f.message = "the first line is not \"# OOMMF OVF 2.0\"";
        return f;
    }
    f.version = 2;
    const long cp = count_pos( head );
    if( cp < 0 )
    {
        f.message = "no segment count in the file header";
        return f;
    }
    f.n_segments   = std::atoi( head.c_str() + cp );
    f.count_offset = cp;
    f.is_ovf       = true;
    return f;
}

void write_segment( const std::string & file, const Segment & seg, const double * data, int format, bool append )
{
    const std::string body = segment_text( seg, data, format );
    const FileInfo info    = append ? inspect( file ) : FileInfo();
    if( append && info.found )
    {
        if( !info.is_ovf )
            throw std::runtime_error( "OVF: cannot append to non-OVF file \"" + file + "\"" );
        std::ofstream out( file, std::ios::binary | std::ios::app );
        if( !out )
            throw std::runtime_error( "OVF: cannot open \"" + file + "\" for appending" );
        out << body;
        out.close();
        // increment the zero-padded count in place
        std::fstream io( file, std::ios::binary | std::ios::in | std::ios::out );
        std::string n = std::to_string( info.n_segments + 1 );
        n             = std::string( (size_t)std::max( 0, COUNT_DIGITS - (int)n.size() ), '0' ) + n;
        io.seekp( info.count_offset );
        io << n;
        return;
    }
    std::ofstream out( file, std::ios::binary | std::ios::trunc );
    if( !out )
        throw std::runtime_error( "OVF: cannot open \"" + file + "\" for writing" );
    std::string head = "# OOMMF OVF 2.0\n#\n# Segment count: ";
    std::string n    = "1";
    head += std::string( (size_t)( COUNT_DIGITS - 1 ), '0' ) + n + "\n";
    out << head << body;
}

// ---------------------------------------------------------------------------
// reading
// ---------------------------------------------------------------------------
static bool starts_with_ci( const std::string & s, size_t pos, const std::string & key )
{
    if( pos + key.size() > s.size() )
        return false;
    for( size_t i = 0; i < key.size(); ++i )
        if( std::tolower( (unsigned char)s[pos + i] ) != std::tolower( (unsigned char)key[i] ) )
            return false;
    return true;
}

// the line starting at pos (without its newline) and the position after it
static std::string line_at( const std::string & s, size_t pos, size_t & next )
{
    size_t e = s.find( '\n', pos );
    if( e == std::string::npos )
        e = s.size();
    next = e < s.size() ? e + 1 : e;
    return s.substr( pos, e - pos );
}

void read_segment( const std::string & file, int idx, Segment & seg, std::vector<double> & data )
{
    bool found             = false;
    const std::string text = read_all( file, found );
    if( !found )
        throw std::runtime_error( "OVF: cannot open \"" + file + "\"" );
    // find the idx-th "# Begin: Segment"
    size_t pos = 0;
    int k      = -1;
    while( pos < text.size() )
    {
        size_t next;
        std::string ln = line_at( text, pos, next );
        std::string t  = lower( trim( ln ) );
        if( t == "# begin: segment" && ++k == idx )
        {
            pos = next;
            break;
        }
        pos = next;
    }
    if( k != idx )
        throw std::runtime_error( "OVF: file \"" + file + "\" has no segment " + std::to_string( idx ) );
    seg          = Segment();
    seg.meshtype = "rectangular";
    // header lines "# key: value", "##" starts a comment
    std::string dtype;
    while( pos < text.size() )
    {
        size_t next;
        std::string ln = line_at( text, pos, next );
        pos            = next;
        std::string t  = trim( ln );
        if( t.size() < 1 || t[0] != '#' )
            continue;
        if( t.compare( 0, 2, "##" ) == 0 )
            continue;
        std::string body = t.substr( 1 );
        size_t cc        = body.find( "##" );
        if( cc != std::string::npos )
            body = body.substr( 0, cc );
        size_t colon = body.find( ':' );
        if( colon == std::string::npos )
            continue;
        const std::string key = lower( trim( body.substr( 0, colon ) ) );
        const std::string val = trim( body.substr( colon + 1 ) );
        if( key == "begin" && starts_with_ci( val, 0, "data" ) )
        {
            dtype = lower( trim( val.substr( 4 ) ) );
            break;
        }
        if( key == "title" )
            seg.title = val;
        else if( key == "desc" )
            seg.comment = val;
        else if( key == "valuedim" )
            seg.valuedim = std::atoi( val.c_str() );
        else if( key == "valueunits" )
            seg.valueunits = val;
        else if( key == "valuelabels" )
            seg.valuelabels = val;
        else if( key == "meshunit" )
            seg.meshunit = val;
        else if( key == "meshtype" )
            seg.meshtype = lower( val );
        else if( key == "pointcount" )
            seg.pointcount = std::atol( val.c_str() );
        else if( key.size() >= 4 && ( key[0] == 'x' || key[0] == 'y' || key[0] == 'z' ) )
        {
            const int a           = key[0] - 'x';
            const std::string rest = key.substr( 1 );
            const double v        = std::strtod( val.c_str(), nullptr );
            if( rest == "min" )
                seg.bounds_min[a] = v;
            else if( rest == "max" )
                seg.bounds_max[a] = v;
            else if( rest == "base" )
                seg.origin[a] = v;
            else if( rest == "stepsize" )
                seg.step_size[a] = v;
            else if( rest == "nodes" )
                seg.n_cells[a] = std::atoi( val.c_str() );
        }
    }
    if( dtype.empty() )
        throw std::runtime_error( "OVF: segment " + std::to_string( idx ) + " of \"" + file + "\" has no data block" );
    const long rows = seg.rows();
    const long n    = rows * seg.valuedim;
    if( n <= 0 )
        throw std::runtime_error( "OVF: segment " + std::to_string( idx ) + " of \"" + file + "\" holds no values" );
    data.assign( (size_t)n, 0.0 );
    if( dtype == "binary 8" || dtype == "binary 4" )
    {
        const size_t w = dtype == "binary 8" ? 8 : 4;
        if( pos + w * ( (size_t)n + 1 ) > text.size() )
            throw std::runtime_error( "OVF: data block of segment " + std::to_string( idx ) + " is truncated" );
        if( w == 8 )
        {
            uint64_t chk;
            std::memcpy( &chk, text.data() + pos, 8 );
            if( chk != CHECK_8B )
                throw std::runtime_error( "OVF: wrong check value in a Binary 8 block" );
            std::memcpy( data.data(), text.data() + pos + 8, 8 * (size_t)n );
        }
        else
        {
            uint32_t chk;
            std::memcpy( &chk, text.data() + pos, 4 );
            if( chk != CHECK_4B )
                throw std::runtime_error( "OVF: wrong check value in a Binary 4 block" );
            std::vector<float> f( (size_t)n );
            std::memcpy( f.data(), text.data() + pos + 4, 4 * (size_t)n );
            for( long i = 0; i < n; ++i )
                data[i] = f[i];
        }
        return;
    }
    if( dtype != "text" && dtype != "csv" )
        throw std::runtime_error( "OVF: unknown data type \"" + dtype + "\"" );
    long got = 0;
    while( pos < text.size() && got < n )
    {
        size_t next;
        std::string ln = line_at( text, pos, next );
        pos            = next;
        std::string t  = trim( ln );
        if( t.empty() )
            continue;
        if( t[0] == '#' )
        {
            if( starts_with_ci( t, 0, "# end: data" ) )
                break;
            continue;
        }
        const char * c = t.c_str();
        while( *c && got < n )
        {
            while( *c == ' ' || *c == '\t' || *c == ',' )
                ++c;
            if( !*c )
                break;
            char * end;
            double v = std::strtod( c, &end );
            if( end == c )
                throw std::runtime_error( "OVF: unreadable value in a text data block" );
            data[got++] = v;
            c           = end;
        }
    }
    if( got < n )
        throw std::runtime_error( "OVF: text data block of segment " + std::to_string( idx ) + " holds "
                                  + std::to_string( got ) + " of " + std::to_string( n ) + " values" );
}

} // namespace ovf
} // namespace spirit_hip
