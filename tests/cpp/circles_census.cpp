// Circle census of a complete surface, driven the way app/circles/main.cu:13-87 drives the reference:
// getData() plus the internal getters copied to the host with DeviceBuffer::toHost<Type>(). Circles
// (i < j) are counted as buried (status 0), full (2), intersected without a visible intersection
// (1) and intersected with segments (marked 3 here, as the reference does). Every segment's circle
// must have status 1 ("Found non intersecting circle" in the reference): such findings are counted.
// Reads float32 xyzr atoms from a file; prints one JSON line.
#include <cstdio>
#include <optional>
#include <vector>

#include "bcs/sesdf/sesdf.hpp"

struct Census
{
    uint32_t buried = 0, full = 0, withoutIntersections = 0, withIntersections = 0, total = 0;
    uint32_t segmentNotIntersected = 0, unknownNeighbor = 0;
};

static Census getCircleData( bcs::Sesdf & sesdf )
{
    bcs::sesdf::SesdfData data = sesdf.getData();

    const auto neighborCount   = sesdf.getDAtomNeighborsCount().toHost<uint32_t>();
    const auto neighborIndices = sesdf.getDAtomNeighborsIndices().toHost<uint32_t>();
    auto       circlesStatus   = sesdf.getDCircleVisibilityStatus().toHost<uint8_t>();
    const auto segments        = bcs::DeviceBuffer( data.segmentPatches, sizeof( uint4 ) * data.segmentPatchNb ).toHost<uint4>();

    Census         out;
    const uint32_t maxNeighborsPerAtom = sesdf.getMaxNeighborPerAtom();
    for ( const uint4 segment : segments )
    {
        const uint32_t          neighborNb = neighborCount[ segment.x ];
        std::optional<uint32_t> localJ;
        for ( uint32_t l = 0; !localJ && l < neighborNb; l++ )
            if ( neighborIndices[ segment.x * maxNeighborsPerAtom + l ] == segment.y )
                localJ = l;
        if ( !localJ )
        {
            out.unknownNeighbor++;
            continue;
        }
        uint8_t & current = circlesStatus[ segment.x * maxNeighborsPerAtom + *localJ ];
        if ( current != 1 && current != 3 )
            out.segmentNotIntersected++;
        current = 3;
    }
    for ( uint32_t i = 0; i < data.atomNb; i++ )
        for ( uint32_t l = 0; l < neighborCount[ i ]; l++ )
        {
            const uint32_t j = neighborIndices[ i * maxNeighborsPerAtom + l ];
            if ( j < i )
                continue; // each circle once
            out.total++;
            switch ( circlesStatus[ i * maxNeighborsPerAtom + l ] )
            {
            case 0: out.buried++; break;
            case 1: out.withoutIntersections++; break;
            case 2: out.full++; break;
            case 3: out.withIntersections++; break;
            default: break;
            }
        }
    return out;
}

int main( int argc, char ** argv )
{
    if ( argc < 2 )
    {
        std::fprintf( stderr, "usage: circles_census atoms.f32\n" );
        return 2;
    }
    std::FILE * f = std::fopen( argv[ 1 ], "rb" );
    if ( !f )
        return 2;
    std::vector<bcs::Vec4f> molecule;
    bcs::Vec4f              a;
    while ( std::fread( &a, sizeof( a ), 1, f ) == 1 )
        molecule.push_back( a );
    std::fclose( f );

    bcs::Sesdf   sesdf { molecule, bcs::getAabb( molecule ), 1.4f, true };
    const Census c = getCircleData( sesdf );
    // the reference's globalToTrimmed map: its last entry counts the intersected circles
    const auto g2t = sesdf.getDGlobalToTrimmedId().toHost<uint32_t>();
    std::printf( "{\"atoms\": %zu, \"buried\": %u, \"full\": %u, \"without_intersections\": %u, \"with_intersections\": %u, "
                 "\"total\": %u, \"segment_not_intersected\": %u, \"unknown_neighbor\": %u, \"g2t_size\": %zu, \"g2t_last\": %u, "
                 "\"intersected\": %u, \"status\": %d}\n",
                 molecule.size(), c.buried, c.full, c.withoutIntersections, c.withIntersections, c.total, c.segmentNotIntersected,
                 c.unknownNeighbor, g2t.size(), g2t.empty() ? 0u : g2t.back(), sesdf.getCounts().intersected_circle_nb,
                 int( sesdf.lastStatus() ) );
    return 0;
}
