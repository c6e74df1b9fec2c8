// EDI (SEG Electrical Data Interchange) reader / writer for MT data, one
// station per file: the reference's io::read_edi / write_edi /
// load_edi_directory / save_edi_directory (include/maple3dmt/io/edi_io.h,
// src/io/edi_io.cpp:156-619) over this build's data::MTData.
//
// Impedances are stored in EDI units (mV/km/nT) and held in SI (ohm):
// Z_SI = Z_EDI * mu0 * 1e3 = Z_EDI * 4 pi 1e-4; errors are the square root of
// the EDI variance in the same units, 1 * 4 pi 1e-4 when a variance is absent
// (0.1 for tipper); a component with no section has weight 0.  Re and Im are
// taken as written (no conjugation).
#pragma once

#include <string>
#include <vector>

#include "mt_data.h"

namespace m3d {
namespace io {

struct EDIRecord {
    data::Station station;
    std::vector<double> frequencies;           // Hz, file order
    std::vector<data::MTResponse> responses;   // per frequency
};

// Throws std::runtime_error when the file cannot be read or has no usable
// >FREQ section.
EDIRecord read_edi(const std::string& path);
void write_edi(const std::string& path, const data::Station& station, const std::vector<double>& frequencies,
               const std::vector<data::MTResponse>& responses);

// Every *.edi (any case) of dir, in file-name order; unreadable files are
// skipped.  Frequencies: the union over stations (equal within 1e-6
// relative), highest first; a station-frequency pair without data gets
// weight 0 on every component.
void load_edi_directory(const std::string& dir, data::MTData& out);
// One file per station, <name>.edi (station_<i>.edi when unnamed).
void save_edi_directory(const std::string& dir, const data::MTData& d);

}  // namespace io
}  // namespace m3d
