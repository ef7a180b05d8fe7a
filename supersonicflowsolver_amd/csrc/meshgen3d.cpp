// meshgen3d.cpp -- 3D extrusion of the wedge profile (hexahedra, or six
// tetrahedra per hexahedron).  Not yet implemented.
#include <stdexcept>

#include "psst_host.hpp"

namespace psst {

RawMesh generate_wedge_3d(const WedgeParams &) {
    throw std::runtime_error("3D wedge meshes are not implemented yet");
}

}  // namespace psst
