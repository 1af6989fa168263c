// Drives the C++ host classes (bcs::Sesdf, bcs::Ssesdf) the way a reference caller does
// (app/benchmarks/main.cu:174-201): atoms + AABB in, build(), getData() / getGraphics(),
// getAtomIndices(). Reads float32 xyzr atoms from a file and prints one JSON line of counts.
#include <cstdio>
#include <vector>

#include "bcs/sesdf/sesdf.hpp"
#include "bcs/ssesdf/ssesdf.hpp"

int main( int argc, char ** argv )
{
    if ( argc < 2 )
    {
        std::fprintf( stderr, "usage: host_api atoms.f32\n" );
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

    const bcs::Aabb aabb = bcs::getAabb( atoms );
    bcs::Sesdf      complete( atoms, aabb, 1.4f );
    complete.build(); // second build on the same span: same results
    const bcs::sesdf::SesdfData     d = complete.getData();
    const bcs::sesdf::SesdfGraphics g = complete.getGraphics();
    const std::vector<uint32_t>     s2i = complete.getAtomIndices();
    uint64_t                        permSum = 0;
    for ( uint32_t v : s2i )
        permSum += v;

    bcs::Ssesdf exterior( atoms, aabb, 1.4f );
    const bcs::ssesdf::SsesdfData e = exterior.getData();

    // moved-from objects stay destructible, moved-to ones keep the surface
    bcs::Sesdf moved = std::move( complete );
    const bool movedOk = moved.getData().concavePatchNb == d.concavePatchNb;

    std::printf( "{\"atoms\": %u, \"sectors\": %u, \"segments\": %u, \"full_circles\": %u, \"concave\": %u, "
                 "\"sector_bytes\": %zu, \"perm_sum\": %llu, \"exterior_torus\": %u, \"exterior_concave\": %u, "
                 "\"max_neighbors\": %u, \"moved\": %s, \"status\": %d}\n",
                 d.atomNb, d.sectorNb, d.segmentPatchNb, d.circlePatchNb, d.concavePatchNb, g.sectors.size,
                 (unsigned long long)permSum, e.circlePatchNb, e.concavePatchNb, unsigned( moved.getMaxNeighborPerAtom() ),
                 movedOk ? "true" : "false", int( moved.lastStatus() ) );
    return 0;
}
