// Host-side state of one builder handle (single molecule or batch).
#pragma once

#include <vector>

#include "components.h"
#include "kernels.h"
#include "sesdf_c.h"

namespace sesdf
{
    struct Capacities
    {
        uint32_t sectors = 0, full = 0, trimmed = 0, records = 0, lists = 0, segments = 0, probeNb = 0, torus = 0;
    };

    // Device state of the SDF evaluator (allocated on first use, refreshed once per build).
    struct SdfState
    {
        Components comp {};                 // surface components (exterior variant), see components.h
        void *     compAlloc  = nullptr;
        uint64_t   compSerial = ~0ull;      // build the component labels belong to
        float4 *   gradScratch = nullptr;  // N: gradient in sorted order (backward without a caller buffer)
        unsigned long long * gradFixed = nullptr; // 4 N + 2: fixed-point accumulators, then the scale words
        uint32_t   gradFixedCap = 0;
        float4 *   packAtoms     = nullptr; // N: centre, R
        float4 *   surfAtoms     = nullptr; // surface atoms, compacted in cell order
        int4 *     surfRanges    = nullptr; // 2 per surface atom: element ranges
        uint32_t * surfIds       = nullptr; // per surface atom: sorted index | face flag << 31
        uint32_t * surfFlag      = nullptr; // N: surface atom (scan input)
        uint32_t * surfInfo      = nullptr; // N: face may be exposed
        uint32_t * surfPrefix    = nullptr; // N + 1: exclusive scan of surfFlag
        uint32_t * surfCellStart = nullptr; // cellNb + 1
        uint32_t * surfTotal     = nullptr;
        int2 *     surfCaps      = nullptr; // per surface atom: range of its visible circles' caps
        float4 *   geo           = nullptr; // caps (cap.sectors) then segment geometry (3 per segment)
        uint32_t   geoSectors    = 0;
        uint32_t   geoSegments   = 0;
        uint64_t   packSerial    = ~0ull; // build the packed / surface arrays belong to
        float4 *   ptSorted    = nullptr;  // point sort space (ptCap): points in cell order, keys and
        uint32_t   ptCap       = 0;        // indices double-buffered, the sort's histograms
        uint64_t * ptCodes     = nullptr;  // backward without stored winners: codes of a forward pass
        float *    ptVals      = nullptr;  //   and its values (ptEvalCap each)
        uint64_t   ptEvalCap   = 0;
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
        bool        released  = false;   // device buffers freed by sesdf_release (the next build allocates)
        bool        growNeighbors = false; // grow K and rebuild when a list overflows (else truncate)
        uint32_t    neighborGrowths = 0;   // K growths by the last build
        bool        inputValid = false;
        uint32_t    rebuilds  = 0;
        uint32_t    relayouts = 0;           // grid lay-outs redone by the last build (moved atoms)
        uint64_t    buildSerial = 0;     // stage sequences enqueued so far (keys the SDF caches)
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

        uint32_t stageNext = 0;          // stage entry points: the next stage allowed (they run in order)
        uint32_t * g2t     = nullptr;    // globalToTrimmed (N * K + 1), computed on request
        uint64_t g2tSerial = ~0ull;

        hipGraphExec_t graphExec = nullptr;
        bool           graphWarm = false; // a build has run on the current buffers: capture from the next one
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
    // Once per build, on the handle's stream. With mayAlloc false, SESDF_ERR_CAPACITY instead of
    // allocating when the component buffers are missing or too small for the build.
    sesdf_status prepare_components( Handle & h, bool mayAlloc = true );
    // Records the message that sesdf_last_error returns; returns st.
    sesdf_status report_error( sesdf_status st, const char * what );
} // namespace sesdf
