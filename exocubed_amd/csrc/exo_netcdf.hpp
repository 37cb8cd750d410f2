// exo_netcdf.hpp -- the reference's per-MeshBlock NetCDF output
// (NetcdfOutput::WriteOutputFile, src/outputs/netcdf.cpp:34-469): file names,
// dimensions, coordinate variables and attributes, written as NetCDF-3
// 64-bit-offset files by a self-contained writer (no libnetcdf needed).
#pragma once

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "exo_setup.hpp"

namespace exo {

enum OutSet { OUT_PRIM = 0, OUT_CONS = 1, OUT_UOV = 2 };

// "prim" / "cons" / "uov" -> OutSet, or -1
int out_set_from_name(const char *variable);

// Variable names of one output set in file order (Athena's LoadOutputData
// order: density, pressure or energy, the vector, passive scalars; the user
// outputs of examples/2023-Chen-exo3/W92.cpp:143-150 for "uov").
std::vector<std::string> out_names(int set, int eos, int ntracer);

// Per-column factors of the "uov" fields: [7][6][nc3][nc2] doubles holding
// lat and lon in degrees, sqrtg and the four factors of the lat-lon wind.
std::vector<double> uov_table(const Setup &s);

// Writes the files of one output: one per MeshBlock, the six panels each
// split into blocks_per_side^2 blocks.  fields: [6][nvar][n1][n2][n3] floats
// (x3 fastest) of the whole panels, n = interior cells or, with ghost_zones,
// cells including three ghosts on each side.  Returns "" or an error.
std::string write_output_files(const Setup &s, const std::string &basename, int set,
                               const std::string &file_id, int file_number, double time,
                               int blocks_per_side, int ghost_zones, const float *fields);

}  // namespace exo
