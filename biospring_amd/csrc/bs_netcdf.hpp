// Spring-network topology from a NetCDF file: the variables and checks of the
// reference's NetCDFReader (src/IO/NetCDFReader.cpp:8-280, buffer sizes from
// src/IO/SpnBuffer.h:10-12), read by a self-contained parser of the NetCDF
// classic formats (CDF-1, the 64-bit-offset CDF-2 and the 64-bit-data CDF-5),
// without the netCDF library.  NetCDF-4 (HDF5) files are refused.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace bs
{

// The topology as the reference's reader hands it to Topology::add_particle /
// add_spring and Topology::to_spring_network (Topology.hpp:245-296).
struct NcTopology
{
    int64_t n = 0;
    std::vector<float> x;                              // n x 3
    std::vector<float> mass, charge, radius, epsilon;  // mass 0 -> 1 (NetCDFReader.cpp:34-38)
    std::vector<float> sasa, transfer, hydro;          // optional variables, 0 when absent
    std::vector<uint8_t> is_static;                    // dynamicstate == 0
    std::vector<int32_t> atom_id, resid;
    std::vector<std::string> name, resname, chain;
    std::vector<uint32_t> s1, s2;                      // springs (file order)
    std::vector<float> l0, k;
    std::vector<std::string> warnings;                 // the reader's logging::warning lines
};

// Reads `path`; returns false with `err` set (the reference's die() message
// where it has one).
bool read_netcdf_topology(const std::string &path, NcTopology &out, std::string &err);

// Writes the dimensions, variables and attributes of NetCDFWriter::writeBinary
// (src/IO/NetCDFWriter.cpp:25-150) in the 64-bit-offset classic layout:
// positions, ids (the particle index), properties, dynamic state, names cut
// to their buffer lengths, springs and the springs per particle.
bool write_netcdf_topology(const std::string &path, const NcTopology &t, std::string &err);

} // namespace bs
