// Host file formats around the GPU path: BABELSIM_MESH 3 meshes, per-rank
// result CSVs and metadata, nonuniform cell values for restarts.
#pragma once

#include "mesh_host.hpp"

#include <string>
#include <vector>

namespace bsim {

// src/io/mesh_reader.cpp:44-111 (patch kinds as in kPatch*).
MeshDesc readMeshFile(const std::string& path);
void writeMeshFile(const MeshDesc& d, const std::string& path);
// src/parallel/parallel_writer.cpp:42-75: <timeDir>/rank-NNNN/<name>.csv.
void writeFieldCsv(const std::string& timeDir, int rank, const std::string& name, int ncells, const int* globalIds,
                   const double* centres, int ncomp, const double* values);
// src/parallel/parallel_writer.cpp:116-169: metadata.bs and mesh.geometry.
void writeResultMetadata(const std::string& timeDir, int rank, int ranks, const MeshDesc& d, const int* globalIds,
                         long long globalCells, const std::string& timeName, const std::vector<std::string>& names,
                         const std::vector<std::string>& types, const std::vector<int>& locations);
// src/io/field_reader.cpp:143-241: rank-sharded CSV directory, one CSV, or
// "id value..." rows; every cell exactly once.
void readFieldValues(const std::string& path, const std::string& name, int ncells, int ncomp, double* out);

}  // namespace bsim
