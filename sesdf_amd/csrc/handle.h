// Host-side state of one builder handle (single molecule or batch).
#pragma once

#include <vector>

#include "kernels.h"
#include "sesdf_c.h"

namespace sesdf
{
    struct Capacities
    {
        uint32_t sectors = 0, full = 0, trimmed = 0, records = 0, lists = 0, segments = 0, probeNb = 0, torus = 0;
    };

    // Device state of the SDF evaluator: trimmed SAS boundary elements binned on the atom grid.
    struct SdfState
    {
        bool       valid       = false;
        uint32_t   builtEpoch  = 0;
        uint32_t * exposedMask = nullptr;  // N * K/32: neighbour slots whose cap borders an exposed face
        uint32_t * arcStart    = nullptr;  // N + 1: arcs (full circles + segments) owned by atom i
        float4 *   arcs        = nullptr;  // per arc: 3 float4 (centre+radius, normal+owner, endpoints A/B as angles)
        uint32_t   arcCap      = 0;
        uint32_t * vtxStart    = nullptr;  // N + 1: concave vertices owned by atom i
        uint32_t * component   = nullptr;  // N: surface component label (exterior classification)
        uint32_t * winner      = nullptr;  // per stored query: winning element code
        uint64_t   winnerCap   = 0;
        float4 *   molGrids    = nullptr;  // batches: per-molecule grid origin / spacing
        uint32_t   molGridDim  = 0;
        float      molGridMargin = 0.f;
    };

    struct Handle
    {
        int         device    = 0;
        hipStream_t stream    = nullptr;
        bool        ownStream = false;
        int         mode      = 0;
        uint32_t    n         = 0;
        uint32_t    K         = 128;
        uint32_t    maxProbeNb = 32;
        float       rp        = 1.4f;
        float       cellSize  = 3.2f;
        bool        useGraph  = false;
        bool        allocated = false;
        bool        built     = false;
        bool        inputValid = false;
        uint32_t    rebuilds  = 0;
        uint32_t    stageStride = 32;    // staged concave records per atom (grows on overflow)
        uint32_t    wavesCircles = 4;    // block shapes (autotuned)
        uint32_t    wavesInter   = 4;

        const float *         hostAtoms = nullptr;
        std::vector<Molecule> mols;
        std::vector<float>    boxes;     // 6 per molecule
        uint32_t              cellNb = 0;
        Capacities            cap;

        std::vector<void *> allocs;
        uint64_t            deviceBytes = 0;
        Buffers             b {};
        float4 *            input = nullptr;
        Control *           hctl  = nullptr;  // pinned host copy of the control block

        hipGraphExec_t graphExec = nullptr;
        hipEvent_t     events[ SESDF_STAGE_COUNT + 1 ] = {};

        SdfState            sdf;
        std::vector<void *> sdfAllocs;
        uint64_t            sdfBytes = 0;

        void freeAll();
        void freeSdf()
        {
            for ( void * p : sdfAllocs )
                (void)hipFree( p );
            sdfAllocs.clear();
            sdfBytes = 0;
            sdf      = SdfState {};
        }
    };

    sesdf_status allocate( Handle & h );
    sesdf_status build_resident( Handle & h );
} // namespace sesdf
