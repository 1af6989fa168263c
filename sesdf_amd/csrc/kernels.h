// Kernel arguments and host launch wrappers of the SES stage sequence.
#pragma once

#include "common.h"

namespace sesdf
{
    // Intersected circles per atom whose touch masks are kept (more fall back to testing all of N(i)).
    constexpr uint32_t TouchStride = 32;

    // Every device buffer of one build (a handle owns them in one arena). Passed by value to
    // every kernel; counts that only the device knows are read from `ctl`.
    struct Buffers
    {
        Control *        ctl;
        const Molecule * mols;
        uint32_t         molNb;

        // input and grid
        const float4 * input;          // N input atoms (input order)
        uint32_t *     atomCell;       // N: global cell index of each input atom
        uint32_t *     atomMol;        // N: molecule of each atom (batches; input and sorted ranges coincide)
        uint32_t *     cellCount;      // cellNb: atoms per cell (self-clearing)
        uint32_t *     cellStart;      // cellNb + 1: exclusive scan of cellCount
        uint32_t *     slotInput;      // N: input index placed at each sorted slot (before in-cell sort)
        uint32_t       cellNb;

        // sorted atoms and neighbourhoods
        float4 *   atoms;              // N sorted atoms (xyz, vdW radius)
        uint32_t * sortedToInitial;    // N
        uint32_t * nbCount;            // N
        uint32_t * nbIds;              // N * K (sorted indices, ascending)
        uint8_t *  status;             // N * K (0 hidden, 1 intersected, 2 full)
        uint32_t * intMask;            // N * K/32: bit set for slots with status 1 and j > i

        // per-atom counts of the current stage (4 arrays of N) and their exclusive prefixes
        uint32_t * atomCnt;            // 4 * N
        uint32_t * vStart;             // N + 1: first sector of atom i
        uint32_t * trimBase;           // N + 1: first intersected-circle (trimmed) id of atom i
        uint32_t * fullBase;           // N + 1: first full circle of atom i
        uint32_t * wStart;             // N + 1: first with-neighbour concave record of atom i
        uint32_t * nStart;             // N + 1: first without-neighbour concave record of atom i
        uint32_t * torusBase;          // N + 1: first torus record of atom i
        uint32_t * stage;              // N * stageStride: packed concave records per atom
        uint32_t   stageStride;
        uint4 *    touch;              // N * TouchStride: per intersected circle (i, j > i), in slot order,
                                       // the bit set of neighbour slots whose SAS sphere touches the circle

        // complete-surface outputs
        uint2 *    convexPatches;      // N: [start, end) sector range
        float4 *   sectors;            // capSectors
        uint2 *    fullCircles;        // capFull
        uint32_t * trimmedToGlobal;    // capTrimmed
        uint32_t * circleIntNb;        // capTrimmed: intersections per intersected circle
        uint32_t * circleStart;        // capTrimmed + 1
        uint32_t * segStart;           // capTrimmed + 1
        uint32_t * circleCursor;       // capTrimmed
        uint32_t * circleLists;        // capLists (3 I)
        uint4 *    segments;           // capSegments

        // concave patches
        uint4 *    records;            // capRecords: with-neighbour records grow up, others grow down
        int4 *     concaveIds;         // capRecords
        float4 *   concavePos;         // capRecords
        float4 *   concaveNb;          // capProbeNb * maxIntersectionNeighbors

        // exterior torus records
        float4 * torusPos;             // capTorus
        float4 * torusAxis;            // capTorus
        float4 * torusVs;              // capTorus

        uint64_t * flags;              // look-back flags: 4 arrays of flagStride words
        uint32_t   flagStride;

        uint32_t n;                    // atoms
        uint32_t wavesCircles;         // waves per workgroup of the circle kernel (2, 4, 8)
        uint32_t wavesInter;           // waves per workgroup of the intersection kernel (2, 4, 8)
        uint32_t K;                    // neighbour capacity per atom (multiple of 32)
        uint32_t maxProbeNb;           // probe-neighbour capacity per concave patch (reference: 32)
        float    probeRadius;
        int      mode;                 // 0 complete, 1 exterior path
        uint32_t capSectors, capFull, capTrimmed, capLists, capSegments, capRecords, capProbeNb, capTorus;
    };

    // Exclusive scan of up to 4 per-item count arrays (tiles of ScanMinTile items or more, decoupled
    // look-back on one ticket per tile). out[c] receives n + 1 values (total last); totals[c], when set, too.
    constexpr uint32_t ScanMinTile = 2048; // smallest scan tile: look-back flag arrays hold n / ScanMinTile + 2 words
    struct ScanArgs
    {
        uint32_t *       in[ 4 ];
        uint32_t *       out[ 4 ];
        uint32_t *       totals[ 4 ];
        const uint32_t * countPtr;     // when set, n = min(*countPtr, n)
        uint32_t         n;
        uint32_t         ticket;       // Ticket slot
        uint32_t         stage;        // look-back tag
        int              clearInput;   // zero the inputs after reading (self-clearing counters)
        int              halfSecond;   // out[1] scans in[0] / 2 instead of in[1]
        uint32_t *       clearAlso;    // zeroed alongside (same length), may be null
        uint32_t *       maxOut;       // max of in[0] over the items (atomicMax once per tile)
        uint32_t *       sumOut;       // sum of in[3] (atomicAdd once per tile), scanned arrays < 3
    };
    template<int NC>
    void launch_scan( const Buffers & b, const ScanArgs & a, hipStream_t s );

    // Host launch wrappers (one per stage); all are stream-ordered and capturable.
    void launch_begin( const Buffers & b, hipStream_t s );
    void launch_grid( const Buffers & b, hipStream_t s );
    void launch_circles( const Buffers & b, hipStream_t s );        // neighbour query + visibility
    void launch_circles_finish( const Buffers & b, hipStream_t s ); // offsets + sectors / circle lists
    void launch_intersections( const Buffers & b, hipStream_t s );  // probe candidates + occlusion
    void launch_intersections_finish( const Buffers & b, hipStream_t s ); // offsets + record placement
    void launch_circle_scan( const Buffers & b, hipStream_t s );
    void launch_fill( const Buffers & b, hipStream_t s );
    void launch_segments( const Buffers & b, hipStream_t s );
    void launch_probe_neighbors( const Buffers & b, hipStream_t s );
    void launch_torus( const Buffers & b, hipStream_t s );
    void launch_global_to_trimmed( const Buffers & b, uint32_t * out, hipStream_t s ); // N * K + 1 prefix counts

    int device_cu_count();
} // namespace sesdf
