// The atom container the builders take, as bcs::Molecule (sesdf/include/bcs/core/molecule.hpp:11):
// one xyz + van der Waals radius per atom. Radii by element come from sesdf_element_radius (the
// table of sesdf/src/core/molecule.cpp:9-131).
#ifndef BCS_CORE_MOLECULE_HPP
#define BCS_CORE_MOLECULE_HPP

#include <string>
#include <vector>

#include "bcs/core/type.hpp"
#include "sesdf_c.h"

namespace bcs
{
    using Molecule = std::vector<Vec4f>;

    // Radius of an element symbol (any case); unknown symbols give 1.20 A.
    inline float getRadius( const std::string & element ) { return sesdf_element_radius( element.c_str() ); }
} // namespace bcs

#endif // BCS_CORE_MOLECULE_HPP
