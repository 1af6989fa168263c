// Stage entry points of the exterior-path build (the reference's bcs::ssesdf::findCircles,
// findIntersections, handleIntersectionSingularities and fillCircles,
// sesdf/include/bcs/ssesdf/operations.cuh:11-34), acting on an Ssesdf; see bcs/sesdf/operations.hpp.
#ifndef BCS_SSESDF_OPERATIONS_HPP
#define BCS_SSESDF_OPERATIONS_HPP

#include "bcs/ssesdf/ssesdf.hpp"

namespace bcs::ssesdf
{
    inline sesdf_status upload( Ssesdf & s ) { return sesdf_upload( s.handle() ); }
    inline sesdf_status findCircles( Ssesdf & s ) { return sesdf_find_circles( s.handle() ); }
    inline sesdf_status findIntersections( Ssesdf & s ) { return sesdf_find_intersections( s.handle() ); }
    inline sesdf_status handleIntersectionSingularities( Ssesdf & s ) { return sesdf_handle_intersection_singularities( s.handle() ); }
    inline sesdf_status fillCircles( Ssesdf & s ) { return sesdf_fill_circles( s.handle() ); }
    inline sesdf_status finish( Ssesdf & s ) { return sesdf_sync( s.handle() ); }
} // namespace bcs::ssesdf

#endif // BCS_SSESDF_OPERATIONS_HPP
