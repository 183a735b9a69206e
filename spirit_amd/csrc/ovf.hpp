// OVF 2.0 vector-field files (OOMMF format) as Spirit writes and reads them: a file header
// with a zero-padded segment count, then segments of a text header and a data block in
// text, CSV, or little-endian binary with a check value (1234567.0 as float32,
// 123456789012345.0 as float64).
//
// Restated reference code (paths relative to the reference checkout):
//   ovf::detail::write (segment layout, check values, appending)
//                              core/thirdparty/ovf/include/detail/write.hpp:70-342,
//                              core/thirdparty/ovf/include/detail/helpers.hpp:12-17
//   IO::OVF_Segment (geometry fields of a Spirit image)
//                              core/src/io/OVF_File.cpp:12-39
#pragma once

#include <string>
#include <vector>

namespace spirit_hip
{
namespace ovf
{

enum Format
{
    FORMAT_BIN  = 0, // binary in the data's own precision (double here)
    FORMAT_BIN4 = 1,
    FORMAT_BIN8 = 2,
    FORMAT_TEXT = 3,
    FORMAT_CSV  = 4,
};

struct Segment
{
    std::string title, comment, valueunits, valuelabels, meshunit = "nm", meshtype = "rectangular";
    int valuedim      = 3;
    int n_cells[3]    = { 0, 0, 0 };
    long pointcount   = 0;
    double bounds_min[3] = { 0, 0, 0 }, bounds_max[3] = { 0, 0, 0 };
    double origin[3] = { 0, 0, 0 }, step_size[3] = { 0, 0, 0 };
    long rows() const; // data rows: xnodes*ynodes*znodes (rectangular) or pointcount (irregular)
};

// Write `seg` with data rows[n][valuedim] (row-major) as a new file (append = false) or as
// one more segment of an existing OVF file (append = true; a missing file is created).
// Throws std::runtime_error on invalid formats, segments or I/O errors.
void write_segment( const std::string & file, const Segment & seg, const double * data, int format, bool append );

struct FileInfo
{
    bool found  = false; // the file exists
    bool is_ovf = false; // it starts with an OVF 2.0 header
    int version = 0;
    int n_segments = 0;  // the header's segment count
    long count_offset = -1; // file offset of the segment count's digits
    std::string message; // why it is not OVF
};
FileInfo inspect( const std::string & file );

// Header and data of segment idx (0-based); data holds rows() x valuedim values.
void read_segment( const std::string & file, int idx, Segment & seg, std::vector<double> & data );

} // namespace ovf
} // namespace spirit_hip
