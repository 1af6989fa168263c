// Runs the build stage by stage and times each stage, the way app/detailed/sesdf_benchmark.inl:13-266
// drives the reference's operations (here with a host clock around each stage plus finish(), since
// the stages are stream-ordered). Checks that the staged build gives the counts of build(). Reads
// float32 xyzr atoms; argv[2] = "exterior" runs the exterior path. Prints one JSON line.
#include <chrono>
#include <cstdio>
#include <cstring>
#include <vector>

#include "bcs/sesdf/operations.hpp"
#include "bcs/ssesdf/operations.hpp"

template<class S, class F>
static double timed( S & s, F && stage )
{
    const auto t0 = std::chrono::steady_clock::now();
    const int  st = int( stage( s ) );
    const int  sy = int( sesdf_sync( s.handle() ) ); // waits for the stage (and reads counts back)
    const auto t1 = std::chrono::steady_clock::now();
    if ( st != 0 || sy != 0 )
        return -1.0;
    return std::chrono::duration<double, std::milli>( t1 - t0 ).count();
}

int main( int argc, char ** argv )
{
    if ( argc < 2 )
        return 2;
    std::FILE * f = std::fopen( argv[ 1 ], "rb" );
    if ( !f )
        return 2;
    std::vector<bcs::Vec4f> molecule;
    bcs::Vec4f              a;
    while ( std::fread( &a, sizeof( a ), 1, f ) == 1 )
        molecule.push_back( a );
    std::fclose( f );
    const bool exterior = argc > 2 && std::strcmp( argv[ 2 ], "exterior" ) == 0;
    const auto aabb     = bcs::getAabb( molecule );

    sesdf_counts whole {}, staged {};
    double       ms[ 4 ] = { 0, 0, 0, 0 };
    int          order   = 0;
    if ( !exterior )
    {
        bcs::Sesdf ref { molecule, aabb, 1.4f, true };
        whole = ref.getCounts();
        bcs::Sesdf s { molecule, aabb, 1.4f, false };
        bcs::sesdf::upload( s );
        ms[ 0 ] = timed( s, bcs::sesdf::findCircles );
        ms[ 1 ] = timed( s, bcs::sesdf::findIntersections );
        ms[ 2 ] = timed( s, bcs::sesdf::buildConicPatches );
        ms[ 3 ] = timed( s, bcs::sesdf::handleIntersectionSingularities );
        staged  = s.getCounts();
        // out of order: the intersections again before the circles of a new sequence
        order = int( bcs::sesdf::findIntersections( s ) );
    }
    else
    {
        bcs::Ssesdf ref { molecule, aabb, 1.4f, true };
        whole = ref.getCounts();
        bcs::Ssesdf s { molecule, aabb, 1.4f, false };
        bcs::ssesdf::upload( s );
        ms[ 0 ] = timed( s, bcs::ssesdf::findCircles );
        ms[ 1 ] = timed( s, bcs::ssesdf::findIntersections );
        ms[ 2 ] = timed( s, bcs::ssesdf::handleIntersectionSingularities );
        ms[ 3 ] = timed( s, bcs::ssesdf::fillCircles );
        staged  = s.getCounts();
        order   = int( sesdf_build_conic_patches( s.handle() ) ); // not a stage of the exterior path
    }
    const bool same = whole.neighbor_nb == staged.neighbor_nb && whole.sector_nb == staged.sector_nb &&
                      whole.full_circle_nb == staged.full_circle_nb && whole.concave_patch_nb == staged.concave_patch_nb &&
                      whole.concave_with_neighbors_nb == staged.concave_with_neighbors_nb && whole.segment_nb == staged.segment_nb &&
                      whole.torus_nb == staged.torus_nb;
    std::printf( "{\"atoms\": %zu, \"exterior\": %s, \"ms\": [%.4f, %.4f, %.4f, %.4f], \"same_counts\": %s, \"concave\": %u, "
                 "\"segments\": %u, \"torus\": %u, \"out_of_order_status\": %d}\n",
                 molecule.size(), exterior ? "true" : "false", ms[ 0 ], ms[ 1 ], ms[ 2 ], ms[ 3 ], same ? "true" : "false",
                 staged.concave_patch_nb, staged.segment_nb, staged.torus_nb, order );
    return 0;
}
