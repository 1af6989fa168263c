// Kernel arguments and host launch wrappers of the SES stage sequence.
#pragma once

#include "common.h"

namespace sesdf
{
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
        uint32_t *     cellCount;      // cellNb: atoms per cell (self-clearing)
        uint32_t *     cellStart;      // cellNb + 1: exclusive scan of cellCount
        uint32_t *     cellFill;       // cellNb: scatter cursor (self-clearing)
        uint32_t *     slotInput;      // N: input index placed at each sorted slot (before in-cell sort)
        uint32_t       cellNb;

        // sorted atoms and neighbourhoods
        float4 *   atoms;              // N sorted atoms (xyz, vdW radius)
        uint32_t * sortedToInitial;    // N
        uint32_t * nbCount;            // N
        uint32_t * nbIds;              // N * K (sorted indices, ascending)
        uint8_t *  status;             // N * K (0 hidden, 1 intersected, 2 full)
        uint32_t * intMask;            // N * K/32: bit set for slots with status 1 and j > i
        uint32_t * trimBase;           // N: first trimmed id of atom i
        uint32_t * fullBase;           // N: first full circle of atom i
        uint32_t * nStart;             // N + 1: first without-neighbour record of each atom
        uint32_t * torusBase;          // N + 1: first torus record of each atom

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
        uint32_t * wStart;             // N + 1: first with-neighbour record of each atom
        int4 *     concaveIds;         // capRecords
        float4 *   concavePos;         // capRecords
        float4 *   concaveNb;          // capProbeNb * maxIntersectionNeighbors

        // exterior torus records
        float4 * torusPos;             // capTorus
        float4 * torusAxis;            // capTorus
        float4 * torusVs;              // capTorus

        uint64_t * flags;              // look-back flags: 3 arrays of flagStride words
        uint32_t   flagStride;

        uint32_t n;                    // atoms
        uint32_t K;                    // neighbour capacity per atom (multiple of 32)
        uint32_t maxProbeNb;           // probe-neighbour capacity per concave patch (reference: 32)
        float    probeRadius;
        int      mode;                 // 0 complete, 1 exterior path
        uint32_t capSectors, capFull, capTrimmed, capLists, capSegments, capRecords, capProbeNb, capTorus;
    };

    // Host launch wrappers (one per stage); all are stream-ordered and capturable.
    void launch_begin( const Buffers & b, hipStream_t s );
    void launch_grid( const Buffers & b, hipStream_t s );
    void launch_circles( const Buffers & b, hipStream_t s );
    void launch_intersections( const Buffers & b, hipStream_t s );
    void launch_circle_scan( const Buffers & b, hipStream_t s );
    void launch_fill( const Buffers & b, hipStream_t s );
    void launch_segments( const Buffers & b, hipStream_t s );
    void launch_probe_neighbors( const Buffers & b, hipStream_t s );
    void launch_torus( const Buffers & b, hipStream_t s );

    int device_cu_count();
} // namespace sesdf
