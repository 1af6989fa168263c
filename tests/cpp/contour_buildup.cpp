// Drives bcs::ContourBuildup the way the reference's benchmark does (app/benchmarks/main.cu:155-170):
// molecule + AABB + probe radius + grid size 16, build(), getGraphics(); then a move and a second
// build on the same span. Reads float32 xyzr atoms from a file and prints one JSON line of counts.
#include <cstdio>
#include <vector>

#include "bcs/contour_buildup/contour_buildup.hpp"

int main( int argc, char ** argv )
{
    if ( argc < 2 )
    {
        std::fprintf( stderr, "usage: contour_buildup atoms.f32\n" );
        return 2;
    }
    std::FILE * f = std::fopen( argv[ 1 ], "rb" );
    if ( !f )
        return 2;
    std::vector<bcs::Vec4f> atoms;
    bcs::Vec4f              a;
    while ( std::fread( &a, sizeof( a ), 1, f ) == 1 )
        atoms.push_back( a );
    std::fclose( f );

    const bcs::Aabb         aabb = bcs::getAabb( atoms );
    constexpr uint32_t      gridSize = 16;
    bcs::ContourBuildup     cb( atoms, aabb, 1.4f, gridSize, false );
    cb.build();
    const bcs::cb::ContourBuildupGraphics g = cb.getGraphics();
    const bcs::cb::ContourBuildupData     d = cb.getData();

    bcs::ContourBuildup moved = std::move( cb );
    moved.build(); // second build on the same span: same counts
    const bcs::cb::ContourBuildupGraphics g2 = moved.getGraphics();

    std::printf( "{\"atoms\": %zu, \"probes\": %zu, \"tori\": %zu, \"data_probes\": %u, \"data_tori\": %u, "
                 "\"probe_neighbors\": %u, \"again_probes\": %zu, \"again_tori\": %zu, \"vao\": %u, \"status\": %d}\n",
                 g.atomNb, g.sphericalTriangleNb, g.torusNb, d.sphericalTriangleNb, d.torusNb, d.probeNeighborNb,
                 g2.sphericalTriangleNb, g2.torusNb, g.atomVao + g.sphericalTriangleVao + g.torusVao + g.singularityTexture,
                 int( moved.lastStatus() ) );
    return 0;
}
