// goal's mesh inputs (serial SCOREC .smb, gmi .dmg, association file) read on
// the host into a HostMesh with goal's node, side and element sets.
#pragma once

#include <map>
#include <string>
#include <vector>

#include "goal_mesh.hpp"

namespace goal {

struct MeshFile {
  HostMesh mesh;                                           // node sets and element-set ids filled in
  std::vector<std::string> elem_set_names;                 // element-set id -> name, in file order
  std::map<std::string, std::vector<int32_t>> side_sets;   // name -> 4 * element + local face
};

// Disc's mesh loading (src/goal_disc.cpp:28-99, 347-396).  order 2 adds a node
// at each mesh edge's midpoint, classified like the edge.  Throws
// std::invalid_argument on anything it cannot read.
MeshFile read_mesh_files(const std::string& geom_file, const std::string& mesh_file, const std::string& assoc_file,
                         int order);

}  // namespace goal
