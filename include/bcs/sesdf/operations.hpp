// Stage entry points of the complete-surface build, for callers that run and time the stages one by
// one (app/detailed/sesdf_benchmark.inl:13-266 drives the reference's bcs::sesdf::findCircles,
// findIntersections, buildConicPatches and handleIntersectionSingularities,
// sesdf/include/bcs/sesdf/operations.cuh:11-29). The reference passes its acceleration grid and
// build context; those are internal here, so each stage acts on a Sesdf (its atoms, grid and
// capacities). The stages are stream-ordered; finish() waits for them and reads the counts back.
// Capacities are not grown by these calls: on overflow finish() reports SESDF_ERR_CAPACITY and a
// build() grows them.
#ifndef BCS_SESDF_OPERATIONS_HPP
#define BCS_SESDF_OPERATIONS_HPP

#include "bcs/sesdf/sesdf.hpp"

namespace bcs::sesdf
{
    // host atoms -> device (#1 of bcs::Sesdf::build, sesdf/src/sesdf/sesdf.cu:211-213)
    inline sesdf_status upload( Sesdf & s ) { return sesdf_upload( s.handle() ); }
    inline sesdf_status findCircles( Sesdf & s ) { return sesdf_find_circles( s.handle() ); }
    inline sesdf_status findIntersections( Sesdf & s ) { return sesdf_find_intersections( s.handle() ); }
    inline sesdf_status buildConicPatches( Sesdf & s ) { return sesdf_build_conic_patches( s.handle() ); }
    inline sesdf_status handleIntersectionSingularities( Sesdf & s ) { return sesdf_handle_intersection_singularities( s.handle() ); }
    inline sesdf_status finish( Sesdf & s ) { return sesdf_sync( s.handle() ); }
} // namespace bcs::sesdf

#endif // BCS_SESDF_OPERATIONS_HPP
