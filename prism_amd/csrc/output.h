// output.h -- VTK XML writers of the output export (host code): unstructured-grid .vtu files
// and ParaView .pvd collections.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace prism_hip {

// One data set of a .vtu file: cell or point data, ncomp components interleaved per entity.
struct VtuArray {
    std::string name;
    bool point{false};
    int ncomp{1};
    std::vector<double> data;
};

// Unstructured grid of vertices `pts` [nv][3] and cells given by their vertex lists
// (off [n_cells + 1], conn); the VTK cell type follows the vertex count (8 hexahedron 12,
// 4 tetrahedron 10, 6 wedge 13; anything else throws).  mode 0: ASCII, each double in its
// shortest round-trip decimal form; mode 1: appended raw data, each array zlib-compressed in
// 32 KiB blocks behind a UInt64 block header (vtkZLibDataCompressor).  Point data sets are
// written before cell data sets, each group in the order given.
void writeVtu(const std::string& path, int mode, const std::vector<double>& pts, const std::vector<int64_t>& off,
              const std::vector<int64_t>& conn, const std::vector<VtuArray>& arrays);

// ParaView collection: one DataSet per (time, file), the time printed as "%.6g".
void writePvd(const std::string& path, int64_t n, const double* times, const char* const* files);

}  // namespace prism_hip
