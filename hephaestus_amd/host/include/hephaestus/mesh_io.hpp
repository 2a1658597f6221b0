// MFEM mesh v1.0 text format for the structured hexahedral meshes this path
// runs on (the format of the reference's data/*.mesh files, read by MFEM's
// Mesh constructor in the reference).  ReadMFEMMesh accepts a hexahedral mesh
// whose elements and vertices form a logically structured nx x ny x nz grid
// with the element order and local vertex order of a Cartesian MFEM mesh
// (elements x fastest; v0..v3 the bottom face counter-clockwise, v4..v7 above
// them); element attributes are kept.  Unstructured meshes (tetrahedra, or hexes
// not in that layout) are refused with an Error.
#pragma once

#include "core.hpp"

#include <string>

namespace hephaestus
{

StructuredHexMesh ReadMFEMMesh(const std::string & path);
// Writes mesh in the same layout: element attributes, boundary quads with
// attributes 1..6 for the faces x-, x+, y-, y+, z-, z+, vertices x fastest.
void WriteMFEMMesh(const StructuredHexMesh & mesh, const std::string & path);

} // namespace hephaestus
