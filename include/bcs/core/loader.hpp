// bcs::load (sesdf/include/bcs/core/loader.hpp:8-11, sesdf/src/core/loader.cpp:14-71): the atoms of
// the first model of a structure file, each with its element's radius. MMTF files only (the
// reference reads any format chemfiles knows); other and unreadable files throw std::runtime_error.
#ifndef BCS_CORE_LOADER_HPP
#define BCS_CORE_LOADER_HPP

#include <filesystem>
#include <stdexcept>
#include <string>

#include "bcs/core/molecule.hpp"
#include "sesdf_c.h"

namespace bcs
{
    using Path = std::filesystem::path;

    inline Molecule load( const Path & path )
    {
        float *      xyzr = nullptr;
        uint64_t     n    = 0;
        sesdf_status st   = sesdf_load_mmtf( path.string().c_str(), &xyzr, &n );
        if ( st != SESDF_OK )
            throw std::runtime_error( sesdf_last_error() );
        Molecule molecule( n );
        for ( uint64_t i = 0; i < n; i++ )
            molecule[ i ] = Vec4f { xyzr[ 4 * i ], xyzr[ 4 * i + 1 ], xyzr[ 4 * i + 2 ], xyzr[ 4 * i + 3 ] };
        sesdf_free_atoms( xyzr );
        return molecule;
    }
} // namespace bcs

#endif // BCS_CORE_LOADER_HPP
