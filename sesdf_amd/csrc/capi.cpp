// C ABI of the SES builder: handle lifetime, device arena, capacity growth, hipGraph replay.
//
// The handle plays the part of the reference's host utility classes bcs::Sesdf / bcs::Ssesdf
// (sesdf/src/sesdf/sesdf.cu:20-356, sesdf/src/ssesdf/ssesdf.cu:19-245): it owns every device
// buffer, runs the stage sequence on one stream and exposes the results by device pointer.
// Differences by design: one readback of all counts per build (the reference blocks ~9 times),
// outputs sized by capacities that grow on overflow (then the build is repeated), an explicit
// device instead of "max GFLOPS" selection, and no OpenGL interop.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "handle.h"
#include "sesdf_c.h"

using namespace sesdf;

namespace
{
    thread_local std::string g_lastError;

    sesdf_status hip_fail( hipError_t e, const char * what )
    {
        g_lastError = std::string( what ) + ": " + hipGetErrorString( e );
        return SESDF_ERR_HIP;
    }

    sesdf_status set_error( sesdf_status st, const char * what )
    {
        g_lastError = what;
        return st;
    }

    // Sorted atom indices share a 32-bit word with two flags in the sampling kernels.
    constexpr uint64_t MaxAtoms = 1ull << 30;
    // Largest neighbour capacity K (the kernels' widest list instance).
    constexpr uint32_t MaxNeighbors = 320;
} // namespace

sesdf_status sesdf::report_error( sesdf_status st, const char * what ) { return set_error( st, what ); }

#define HIP_TRY( expr )                                                                                                          \
    do                                                                                                                           \
    {                                                                                                                            \
        hipError_t _e = ( expr );                                                                                                \
        if ( _e != hipSuccess )                                                                                                  \
            return hip_fail( _e, #expr );                                                                                        \
    } while ( 0 )

// Buffers freed by sesdf_release are allocated again before the handle is used.
#define BUFFERS_TRY( h )                                                                                                         \
    do                                                                                                                           \
    {                                                                                                                            \
        if ( ( h ).released )                                                                                                    \
        {                                                                                                                        \
            sesdf_status _s = allocate( h );                                                                                     \
            if ( _s != SESDF_OK )                                                                                                \
                return _s;                                                                                                       \
            ( h ).released = false;                                                                                              \
        }                                                                                                                        \
    } while ( 0 )

// A bounding box the grid can be laid out on: finite, min <= max on each axis (an infinite or
// inverted box would give a brick of no cells, or none that fits).
static bool box_ok( const float * bx )
{
    for ( int d = 0; d < 3; d++ )
        if ( !std::isfinite( bx[ d ] ) || !std::isfinite( bx[ d + 3 ] ) || !( bx[ d ] <= bx[ d + 3 ] ) )
            return false;
    return true;
}

// Every coordinate and radius of atoms [b, e) finite.
static bool atoms_finite( const float * xyzr, uint64_t b, uint64_t e )
{
    for ( uint64_t i = 4 * b; i < 4 * e; i++ )
        if ( !std::isfinite( xyzr[ i ] ) )
            return false;
    return true;
}

static void box_of( const float * xyzr, uint64_t b, uint64_t e, float * out )
{
    float mn[ 3 ] = { 3.4e38f, 3.4e38f, 3.4e38f }, mx[ 3 ] = { -3.4e38f, -3.4e38f, -3.4e38f };
    for ( uint64_t i = b; i < e; i++ )
        for ( int d = 0; d < 3; d++ )
        {
            mn[ d ] = std::min( mn[ d ], xyzr[ 4 * i + d ] - xyzr[ 4 * i + 3 ] );
            mx[ d ] = std::max( mx[ d ], xyzr[ 4 * i + d ] + xyzr[ 4 * i + 3 ] );
        }
    if ( e == b )
        for ( int d = 0; d < 3; d++ )
            mn[ d ] = mx[ d ] = 0.f;
    std::memcpy( out, mn, sizeof( mn ) );
    std::memcpy( out + 3, mx, sizeof( mx ) );
}

namespace sesdf
{
    void Handle::freeAll()
    {
        if ( graphExec )
            (void)hipGraphExecDestroy( graphExec );
        graphExec = nullptr;
        graphWarm = false;
        if ( g2t )
            (void)hipFree( g2t );
        g2t       = nullptr;
        g2tSerial = ~0ull;
        for ( void * p : allocs )
            (void)hipFree( p );
        allocs.clear();
        deviceBytes = 0;
    }

    template<class T>
    static sesdf_status alloc( Handle & h, T *& p, size_t count, bool zero = false )
    {
        const size_t bytes = std::max<size_t>( count, 1 ) * sizeof( T );
        void *       raw   = nullptr;
        HIP_TRY( hipMalloc( &raw, bytes ) );
        if ( zero )
            HIP_TRY( hipMemset( raw, 0, bytes ) );
        h.allocs.push_back( raw );
        h.deviceBytes += bytes;
        p = static_cast<T *>( raw );
        return SESDF_OK;
    }

#define ALLOC_TRY( ... )                                                                                                         \
    do                                                                                                                           \
    {                                                                                                                            \
        sesdf_status _s = alloc( __VA_ARGS__ );                                                                                  \
        if ( _s != SESDF_OK )                                                                                                    \
            return _s;                                                                                                           \
    } while ( 0 )

    // Dense per-molecule cell bricks from each molecule's vdW bounding box. Cell indices are u32:
    // a layout of 2^32 cells or more is refused.
    static sesdf_status layout_grid( Handle & h )
    {
        uint64_t cellBase = 0;
        for ( size_t m = 0; m < h.mols.size(); m++ )
        {
            Molecule & M    = h.mols[ m ];
            float      cell = h.cellSize;
            const float ext[ 3 ] = { h.boxes[ 6 * m + 3 ] - h.boxes[ 6 * m + 0 ], h.boxes[ 6 * m + 4 ] - h.boxes[ 6 * m + 1 ],
                                     h.boxes[ 6 * m + 5 ] - h.boxes[ 6 * m + 2 ] };
            // cap a brick at 2^26 cells by enlarging its cells
            for ( ;; )
            {
                const uint64_t cells = uint64_t( std::floor( ext[ 0 ] / cell ) + 1 ) * uint64_t( std::floor( ext[ 1 ] / cell ) + 1 ) *
                                       uint64_t( std::floor( ext[ 2 ] / cell ) + 1 );
                if ( cells <= ( 1ull << 26 ) )
                    break;
                cell *= 1.25f;
            }
            M.ox       = h.boxes[ 6 * m + 0 ];
            M.oy       = h.boxes[ 6 * m + 1 ];
            M.oz       = h.boxes[ 6 * m + 2 ];
            M.cell     = cell;
            M.dx       = int( std::floor( ext[ 0 ] / cell ) ) + 1;
            M.dy       = int( std::floor( ext[ 1 ] / cell ) ) + 1;
            M.dz       = int( std::floor( ext[ 2 ] / cell ) ) + 1;
            M.cellBase = uint32_t( cellBase );
            cellBase += uint64_t( M.dx ) * M.dy * M.dz;
            if ( cellBase >= 0xfffffff0ull )
                return set_error( SESDF_ERR_CAPACITY, "grid bricks of all molecules exceed 2^32 cells (use a larger cell size)" );
        }
        h.cellNb = uint32_t( cellBase );
        return SESDF_OK;
    }

    static void initial_capacities( Handle & h )
    {
        const uint64_t n = h.n;
        auto           c = []( uint64_t v ) { return uint32_t( std::min<uint64_t>( std::max<uint64_t>( v, 1024 ), 0xfffffff0ull ) ); };
        h.cap.sectors  = h.mode == 0 ? c( 32 * n ) : 1;
        h.cap.full     = h.mode == 0 ? c( n / 2 ) : 1;
        h.cap.trimmed  = h.mode == 0 ? c( 16 * n ) : 1;
        h.cap.records  = c( 2 * n );
        h.cap.lists    = h.mode == 0 ? c( 3ull * h.cap.records ) : 1;
        h.cap.segments = h.mode == 0 ? c( 2ull * h.cap.records ) : 1;
        h.cap.probeNb  = c( h.cap.records / 2 );
        h.cap.torus    = h.mode == 1 ? c( 8 * n ) : 1;
    }

    sesdf_status allocate( Handle & h )
    {
        h.freeAll();
        const uint64_t n = h.n;
        Buffers &      b = h.b;
        b                = Buffers {};
        ALLOC_TRY( h, b.ctl, 1, true );
        Molecule * mols = nullptr;
        ALLOC_TRY( h, mols, h.mols.size() );
        HIP_TRY( hipMemcpy( mols, h.mols.data(), h.mols.size() * sizeof( Molecule ), hipMemcpyHostToDevice ) );
        b.mols  = mols;
        b.molNb = uint32_t( h.mols.size() );
        float4 * input = nullptr;
        ALLOC_TRY( h, input, n );
        h.input = input;
        b.input = input;
        ALLOC_TRY( h, b.atomCell, n );
        ALLOC_TRY( h, b.atomMol, h.mols.size() > 1 ? n : 1 );
        ALLOC_TRY( h, b.cellCount, h.cellNb, true );
        ALLOC_TRY( h, b.cellStart, uint64_t( h.cellNb ) + 1 );
        ALLOC_TRY( h, b.slotInput, n );
        b.cellNb = h.cellNb;
        ALLOC_TRY( h, b.atoms, n );
        ALLOC_TRY( h, b.sortedToInitial, n );
        ALLOC_TRY( h, b.nbCount, n, true );
        ALLOC_TRY( h, b.nbIds, n * h.K );
        ALLOC_TRY( h, b.status, n * h.K + 4, true );
        ALLOC_TRY( h, b.intMask, n * ( h.K / 32 ) );
        ALLOC_TRY( h, b.atomCnt, 4 * n );
        ALLOC_TRY( h, b.vStart, n + 1 );
        ALLOC_TRY( h, b.trimBase, n + 1 );
        ALLOC_TRY( h, b.fullBase, n + 1 );
        ALLOC_TRY( h, b.wStart, n + 1 );
        ALLOC_TRY( h, b.nStart, n + 1 );
        ALLOC_TRY( h, b.torusBase, n + 1 );
        ALLOC_TRY( h, b.stage, n * h.stageStride );
        b.stageStride = h.stageStride;
        ALLOC_TRY( h, b.touch, h.K <= 128 ? n * TouchStride : 1 );
        ALLOC_TRY( h, b.convexPatches, n );
        ALLOC_TRY( h, b.sectors, h.cap.sectors );
        ALLOC_TRY( h, b.fullCircles, h.cap.full );
        ALLOC_TRY( h, b.trimmedToGlobal, h.cap.trimmed );
        ALLOC_TRY( h, b.circleIntNb, h.cap.trimmed );
        ALLOC_TRY( h, b.circleStart, uint64_t( h.cap.trimmed ) + 1 );
        ALLOC_TRY( h, b.segStart, uint64_t( h.cap.trimmed ) + 1 );
        ALLOC_TRY( h, b.circleCursor, h.cap.trimmed );
        ALLOC_TRY( h, b.circleLists, h.cap.lists );
        ALLOC_TRY( h, b.segments, h.cap.segments );
        ALLOC_TRY( h, b.records, h.cap.records );
        ALLOC_TRY( h, b.concaveIds, h.cap.records );
        ALLOC_TRY( h, b.concavePos, h.cap.records );
        ALLOC_TRY( h, b.concaveNb, uint64_t( h.cap.probeNb ) * h.maxProbeNb );
        ALLOC_TRY( h, b.torusPos, h.cap.torus );
        ALLOC_TRY( h, b.torusAxis, h.cap.torus );
        ALLOC_TRY( h, b.torusVs, h.cap.torus );
        // look-back flags: one word per scan tile (ScanMinTile items or more)
        const uint64_t tiles = std::max<uint64_t>( std::max<uint64_t>( h.cellNb, n ), h.cap.trimmed ) / ScanMinTile + 2;
        b.flagStride         = uint32_t( tiles + 64 );
        ALLOC_TRY( h, b.flags, uint64_t( b.flagStride ) * 4, true );

        b.n            = h.n;
        b.wavesCircles = h.wavesCircles;
        b.wavesInter   = h.wavesInter;
        b.K            = h.K;
        b.maxProbeNb  = h.maxProbeNb;
        b.probeRadius = h.rp;
        b.mode        = h.mode;
        b.capSectors  = h.cap.sectors;
        b.capFull     = h.cap.full;
        b.capTrimmed  = h.cap.trimmed;
        b.capLists    = h.cap.lists;
        b.capSegments = h.cap.segments;
        b.capRecords  = h.cap.records;
        b.capProbeNb  = h.cap.probeNb;
        b.capTorus    = h.cap.torus;
        h.allocated   = true;
        h.inputValid  = false;
        return SESDF_OK;
    }

    // The stage sequence of one build (capturable: kernel launches only).
    void enqueue_stages( Handle & h, hipStream_t s, hipEvent_t * ev )
    {
        const Buffers & b = h.b;
        int             e = 0;
        auto            mark = [ & ]() {
            if ( ev )
                (void)hipEventRecord( ev[ e++ ], s );
        };
        mark();
        launch_begin( b, s );
        launch_grid( b, s );
        mark();
        launch_circles( b, s );
        mark();
        launch_circles_finish( b, s );
        mark();
        launch_intersections( b, s );
        mark();
        launch_intersections_finish( b, s );
        mark();
        launch_circle_scan( b, s );
        mark();
        launch_fill( b, s );
        mark();
        launch_segments( b, s );
        mark();
        launch_probe_neighbors( b, s );
        mark();
        launch_torus( b, s );
        mark();
    }

    sesdf_status enqueue( Handle & h )
    {
        HIP_TRY( hipSetDevice( h.device ) );
        h.buildSerial++; // outputs of the previous build are stale from here on
        h.stageNext = 0;
        if ( h.n == 0 )
            return SESDF_OK;
        // the first build on new buffers runs its stages directly: a handle built once (or rebuilt
        // once after sesdf_release) does not pay a graph capture and instantiation it never replays
        if ( h.useGraph && h.graphWarm )
        {
            if ( !h.graphExec )
            {
                hipGraph_t graph = nullptr;
                HIP_TRY( hipStreamBeginCapture( h.stream, hipStreamCaptureModeThreadLocal ) );
                enqueue_stages( h, h.stream, nullptr );
                HIP_TRY( hipStreamEndCapture( h.stream, &graph ) );
                HIP_TRY( hipGraphInstantiate( &h.graphExec, graph, nullptr, nullptr, 0 ) );
                HIP_TRY( hipGraphDestroy( graph ) );
            }
            HIP_TRY( hipGraphLaunch( h.graphExec, h.stream ) );
        }
        else
        {
            enqueue_stages( h, h.stream, nullptr );
            h.graphWarm = true;
        }
        HIP_TRY( hipGetLastError() );
        return SESDF_OK;
    }

    sesdf_status readback( Handle & h )
    {
        if ( h.n == 0 )
        {
            std::memset( h.hctl, 0, sizeof( Control ) );
            return SESDF_OK;
        }
        HIP_TRY( hipMemcpyAsync( h.hctl, h.b.ctl, sizeof( Control ), hipMemcpyDeviceToHost, h.stream ) );
        HIP_TRY( hipStreamSynchronize( h.stream ) );
        return SESDF_OK;
    }

    // Grow every overflowing capacity to 1.25x the reported need; returns true if anything grew.
    static bool grow( Handle & h )
    {
        const Control & c     = *h.hctl;
        bool            grown = false;
        auto            need  = [ & ]( uint32_t & cap, uint64_t want ) {
            want = want + want / 4 + 1024;
            if ( want > cap )
            {
                cap   = uint32_t( std::min<uint64_t>( want, 0xfffffff0ull ) );
                grown = true;
            }
        };
        if ( h.mode == 0 )
        {
            need( h.cap.sectors, c.sectorNb );
            need( h.cap.full, c.fullCircleNb );
            need( h.cap.trimmed, c.intersectedCircleNb );
        }
        const uint64_t I = uint64_t( c.intersectionWithNbNb ) + c.intersectionNoNbNb;
        need( h.cap.records, I );
        need( h.cap.probeNb, c.intersectionWithNbNb );
        if ( c.overflow & OverflowAtomRecords )
        {
            h.stageStride *= 2;
            grown = true;
        }
        if ( h.mode == 0 )
        {
            need( h.cap.lists, 3 * I );
            need( h.cap.segments, std::max<uint64_t>( c.segmentNb, 2 * I ) );
        }
        else
            need( h.cap.torus, c.torusNb );
        return grown;
    }

    static bool overflowed( const Handle & h )
    {
        const Control & c = *h.hctl;
        if ( c.overflow )
            return true;
        const uint64_t I = uint64_t( c.intersectionWithNbNb ) + c.intersectionNoNbNb;
        return I > h.cap.records || c.intersectionWithNbNb > h.cap.probeNb;
    }

    // Neighbour lists longer than K: with the grow policy, K becomes the longest list rounded up to
    // a multiple of 32 (at most 320) and the build runs again on reallocated buffers.
    static sesdf_status grow_neighbors( Handle & h, bool & grown )
    {
        grown = false;
        const Control & c = *h.hctl;
        if ( !h.growNeighbors || c.nbOverflowAtoms == 0 || h.K >= MaxNeighbors )
            return SESDF_OK;
        const uint32_t want = std::min<uint32_t>( MaxNeighbors, ( std::max( c.nbMaxCount, h.K + 1 ) + 31u ) & ~31u );
        std::vector<float4> keep( h.n );
        HIP_TRY( hipMemcpy( keep.data(), h.input, size_t( h.n ) * sizeof( float4 ), hipMemcpyDeviceToHost ) );
        h.K = want;
        h.freeSdf(); // sampling state reads the N x K lists
        const sesdf_status st = allocate( h );
        if ( st != SESDF_OK )
            return st;
        HIP_TRY( hipMemcpy( h.input, keep.data(), size_t( h.n ) * sizeof( float4 ), hipMemcpyHostToDevice ) );
        h.neighborGrowths++;
        grown = true;
        return SESDF_OK;
    }

    sesdf_status build_resident( Handle & h )
    {
        h.rebuilds = 0, h.relayouts = 0, h.neighborGrowths = 0;
        for ( int attempt = 0; attempt < 6; attempt++ )
        {
            sesdf_status st = enqueue( h );
            if ( st != SESDF_OK )
                return st;
            st = readback( h );
            if ( st != SESDF_OK )
                return st;
            bool grownK = false;
            st          = grow_neighbors( h, grownK );
            if ( st != SESDF_OK )
                return st;
            if ( grownK )
                continue;
            if ( !overflowed( h ) )
            {
                h.built = true;
                return SESDF_OK;
            }
            // capacities too small, or atoms outside the grid: grow / lay the grid out again over the
            // current atoms, keep the uploaded input, rebuild
            std::vector<float4> keep( h.n );
            HIP_TRY( hipMemcpy( keep.data(), h.input, size_t( h.n ) * sizeof( float4 ), hipMemcpyDeviceToHost ) );
            const bool relaid = ( h.hctl->overflow & OverflowGrid ) != 0;
            if ( relaid )
            {
                const float * xyzr = reinterpret_cast<const float *>( keep.data() );
                if ( !atoms_finite( xyzr, 0, h.n ) )
                    return set_error( SESDF_ERR_INVALID, "atom coordinates and radii must be finite" );
                for ( size_t m = 0; m < h.mols.size(); m++ )
                    box_of( xyzr, h.mols[ m ].atomBase, uint64_t( h.mols[ m ].atomBase ) + h.mols[ m ].atomNb, h.boxes.data() + 6 * m );
                const sesdf_status lt = layout_grid( h );
                if ( lt != SESDF_OK )
                    return lt;
                h.freeSdf(); // per-cell SDF arrays follow the grid
                h.relayouts++;
            }
            if ( !grow( h ) && !relaid )
            {
                g_lastError = "capacity overflow without growth";
                return SESDF_ERR_CAPACITY;
            }
            st = allocate( h );
            if ( st != SESDF_OK )
                return st;
            HIP_TRY( hipMemcpy( h.input, keep.data(), size_t( h.n ) * sizeof( float4 ), hipMemcpyHostToDevice ) );
            h.rebuilds++;
        }
        g_lastError = "capacity growth did not converge";
        return SESDF_ERR_CAPACITY;
    }
} // namespace sesdf

static sesdf_status common_setup( Handle & h, int device, void * stream, uint32_t K, uint32_t maxProbeNb, float rp, int mode, float cell )
{
    if ( K == 0 )
        K = 128;
    if ( K % 32 != 0 || K > MaxNeighbors || ( mode != 0 && mode != 1 ) || !( rp >= 0.f ) )
        return SESDF_ERR_INVALID;
    h.device     = device;
    h.K          = K;
    h.maxProbeNb = maxProbeNb ? maxProbeNb : 32;
    h.rp         = rp;
    h.mode       = mode;
    h.cellSize   = cell > 0.f ? cell : 3.2f;
    HIP_TRY( hipSetDevice( device ) );
    if ( stream )
        h.stream = static_cast<hipStream_t>( stream );
    else
    {
        HIP_TRY( hipStreamCreateWithFlags( &h.stream, hipStreamNonBlocking ) );
        h.ownStream = true;
    }
    HIP_TRY( hipHostMalloc( reinterpret_cast<void **>( &h.hctl ), sizeof( Control ), hipHostMallocDefault ) );
    std::memset( h.hctl, 0, sizeof( Control ) );
    return SESDF_OK;
}

extern "C"
{
    const char * sesdf_last_error( void ) { return g_lastError.c_str(); }

    const char * sesdf_status_string( sesdf_status s )
    {
        switch ( s )
        {
        case SESDF_OK: return "ok";
        case SESDF_ERR_INVALID: return "invalid argument";
        case SESDF_ERR_HIP: return "HIP runtime error";
        case SESDF_ERR_UNSUPPORTED: return "unsupported";
        case SESDF_ERR_CAPACITY: return "capacity";
        case SESDF_ERR_NOT_BUILT: return "not built";
        }
        return "unknown";
    }

    sesdf_status sesdf_create( const sesdf_create_info * info, sesdf_handle ** out )
    {
        if ( !info || !out )
            return SESDF_ERR_INVALID;
        *out = nullptr;
        if ( info->graphics )
        {
            g_lastError = "graphics (OpenGL interop) is not supported";
            return SESDF_ERR_UNSUPPORTED;
        }
        if ( info->atom_nb && !info->xyzr )
            return SESDF_ERR_INVALID;
        if ( info->atom_nb >= MaxAtoms )
            return set_error( SESDF_ERR_CAPACITY, "at most 2^30 - 1 atoms per handle" );
        // input checks on the host, before any device work
        if ( !atoms_finite( info->xyzr, 0, info->atom_nb ) )
            return set_error( SESDF_ERR_INVALID, "atom coordinates and radii must be finite" );
        {
            bool given = false;
            for ( int d = 0; d < 6; d++ )
                given = given || info->aabb[ d ] != 0.f;
            if ( given && !box_ok( info->aabb ) )
                return set_error( SESDF_ERR_INVALID, "aabb must be finite with min <= max" );
        }
        auto * h = new Handle();
        sesdf_status st = common_setup( *h, info->device, info->stream, info->max_neighbors, info->max_intersection_neighbors,
                                        info->probe_radius, info->mode, info->cell_size );
        if ( st != SESDF_OK )
        {
            delete h;
            return st;
        }
        h->n         = info->atom_nb;
        h->hostAtoms = info->xyzr;
        h->useGraph  = info->use_graph != 0;
        h->boxes.resize( 6 );
        bool haveBox = false;
        for ( int d = 0; d < 6; d++ )
            haveBox = haveBox || info->aabb[ d ] != 0.f;
        if ( haveBox )
            std::memcpy( h->boxes.data(), info->aabb, 6 * sizeof( float ) );
        else
            box_of( info->xyzr, 0, info->atom_nb, h->boxes.data() );
        if ( !box_ok( h->boxes.data() ) )
        {
            sesdf_destroy( reinterpret_cast<sesdf_handle *>( h ) );
            return set_error( SESDF_ERR_INVALID, "aabb (given, or of the atoms) must be finite with min <= max" );
        }
        h->mols.resize( 1 );
        h->mols[ 0 ]          = Molecule {};
        h->mols[ 0 ].atomBase = 0;
        h->mols[ 0 ].atomNb   = h->n;
        st = layout_grid( *h );
        initial_capacities( *h );
        if ( st == SESDF_OK )
            st = allocate( *h );
        if ( st == SESDF_OK && info->build_surface )
            st = sesdf_build( reinterpret_cast<sesdf_handle *>( h ) );
        if ( st != SESDF_OK )
        {
            sesdf_destroy( reinterpret_cast<sesdf_handle *>( h ) );
            return st;
        }
        *out = reinterpret_cast<sesdf_handle *>( h );
        return SESDF_OK;
    }

    sesdf_status sesdf_batch_create( const sesdf_batch_info * info, sesdf_handle ** out )
    {
        if ( !info || !out || !info->atom_offsets || info->molecule_nb == 0 )
            return SESDF_ERR_INVALID;
        *out = nullptr;
        // offsets: molecule m holds atoms [offsets[m], offsets[m + 1]); they start at 0 and never
        // decrease (empty molecules are allowed). Checked before any device work.
        if ( info->atom_offsets[ 0 ] != 0 )
            return set_error( SESDF_ERR_INVALID, "atom_offsets[0] must be 0" );
        for ( uint32_t m = 0; m < info->molecule_nb; m++ )
            if ( info->atom_offsets[ m + 1 ] < info->atom_offsets[ m ] )
                return set_error( SESDF_ERR_INVALID, "atom_offsets must not decrease" );
        if ( info->atom_offsets[ info->molecule_nb ] && !info->xyzr )
            return SESDF_ERR_INVALID;
        if ( info->atom_offsets[ info->molecule_nb ] >= MaxAtoms )
            return set_error( SESDF_ERR_CAPACITY, "at most 2^30 - 1 atoms per handle" );
        if ( info->atom_offsets[ info->molecule_nb ] < 0xffffffffull && !atoms_finite( info->xyzr, 0, info->atom_offsets[ info->molecule_nb ] ) )
            return set_error( SESDF_ERR_INVALID, "atom coordinates and radii must be finite" );
        auto * h = new Handle();
        sesdf_status st = common_setup( *h, info->device, info->stream, info->max_neighbors, info->max_intersection_neighbors,
                                        info->probe_radius, info->mode, info->cell_size );
        if ( st != SESDF_OK )
        {
            delete h;
            return st;
        }
        const uint64_t total = info->atom_offsets[ info->molecule_nb ];
        if ( total >= 0xffffffffull || ( total && !info->xyzr ) )
        {
            sesdf_destroy( reinterpret_cast<sesdf_handle *>( h ) );
            return SESDF_ERR_INVALID;
        }
        h->n         = uint32_t( total );
        h->hostAtoms = info->xyzr;
        h->useGraph  = info->use_graph != 0;
        h->mols.resize( info->molecule_nb );
        h->boxes.resize( 6 * size_t( info->molecule_nb ) );
        for ( uint32_t m = 0; m < info->molecule_nb; m++ )
        {
            h->mols[ m ]          = Molecule {};
            h->mols[ m ].atomBase = uint32_t( info->atom_offsets[ m ] );
            h->mols[ m ].atomNb   = uint32_t( info->atom_offsets[ m + 1 ] - info->atom_offsets[ m ] );
            box_of( info->xyzr, info->atom_offsets[ m ], info->atom_offsets[ m + 1 ], h->boxes.data() + 6 * m );
            if ( !box_ok( h->boxes.data() + 6 * m ) )
            {
                sesdf_destroy( reinterpret_cast<sesdf_handle *>( h ) );
                return set_error( SESDF_ERR_INVALID, "atom coordinates must be finite" );
            }
        }
        st = layout_grid( *h );
        initial_capacities( *h );
        if ( st == SESDF_OK )
            st = allocate( *h );
        if ( st != SESDF_OK )
        {
            sesdf_destroy( reinterpret_cast<sesdf_handle *>( h ) );
            return st;
        }
        *out = reinterpret_cast<sesdf_handle *>( h );
        return SESDF_OK;
    }

    sesdf_status sesdf_destroy( sesdf_handle * hh )
    {
        if ( !hh )
            return SESDF_OK;
        Handle * h = reinterpret_cast<Handle *>( hh );
        (void)hipSetDevice( h->device );
        if ( h->stream )
            (void)hipStreamSynchronize( h->stream );
        h->freeAll();
        h->freeSdf();
        for ( auto & e : h->events )
            if ( e )
                (void)hipEventDestroy( e );
        if ( h->hctl )
            (void)hipHostFree( h->hctl );
        if ( h->ownStream && h->stream )
            (void)hipStreamDestroy( h->stream );
        delete h;
        return SESDF_OK;
    }

    sesdf_status sesdf_set_tuning( sesdf_handle * hh, const sesdf_tuning * t )
    {
        if ( !hh || !t )
            return SESDF_ERR_INVALID;
        Handle & h    = *reinterpret_cast<Handle *>( hh );
        auto     okW  = []( uint32_t w ) { return w == 0 || w == 2 || w == 4 || w == 8; };
        if ( !okW( t->circle_waves ) || !okW( t->intersection_waves ) || t->cell_size < 0.f )
            return SESDF_ERR_INVALID;
        HIP_TRY( hipSetDevice( h.device ) );
        BUFFERS_TRY( h );
        h.wavesCircles = t->circle_waves ? t->circle_waves : 4;
        h.wavesInter   = t->intersection_waves ? t->intersection_waves : 4;
        h.b.wavesCircles = h.wavesCircles;
        h.b.wavesInter   = h.wavesInter;
        if ( h.graphExec )
        {
            HIP_TRY( hipGraphExecDestroy( h.graphExec ) );
            h.graphExec = nullptr;
        }
        if ( t->cell_size > 0.f && t->cell_size != h.cellSize )
        {
            // new grid bricks: re-lay out and reallocate, keeping the uploaded input
            std::vector<float4> keep( h.n );
            if ( h.n )
                HIP_TRY( hipMemcpy( keep.data(), h.input, size_t( h.n ) * sizeof( float4 ), hipMemcpyDeviceToHost ) );
            const float oldCell = h.cellSize;
            h.cellSize          = t->cell_size;
            sesdf_status st     = layout_grid( h );
            if ( st != SESDF_OK )
            {
                h.cellSize = oldCell;
                (void)layout_grid( h );
                return st;
            }
            st = allocate( h );
            if ( st != SESDF_OK )
                return st;
            if ( h.n )
                HIP_TRY( hipMemcpy( h.input, keep.data(), size_t( h.n ) * sizeof( float4 ), hipMemcpyHostToDevice ) );
            h.built = false;
            h.freeSdf();
        }
        return SESDF_OK;
    }

    sesdf_status sesdf_set_atoms( sesdf_handle * hh, const float * xyzr )
    {
        if ( !hh || !xyzr )
            return SESDF_ERR_INVALID;
        reinterpret_cast<Handle *>( hh )->hostAtoms = xyzr;
        return SESDF_OK;
    }

    sesdf_status sesdf_input_buffer( sesdf_handle * hh, float ** d )
    {
        if ( !hh || !d )
            return SESDF_ERR_INVALID;
        Handle & h = *reinterpret_cast<Handle *>( hh );
        HIP_TRY( hipSetDevice( h.device ) );
        BUFFERS_TRY( h );
        *d = reinterpret_cast<float *>( h.input );
        return SESDF_OK;
    }

    sesdf_status sesdf_get_stream( sesdf_handle * hh, void ** s )
    {
        if ( !hh || !s )
            return SESDF_ERR_INVALID;
        *s = reinterpret_cast<Handle *>( hh )->stream;
        return SESDF_OK;
    }

    sesdf_status sesdf_release( sesdf_handle * hh )
    {
        if ( !hh )
            return SESDF_ERR_INVALID;
        Handle & h = *reinterpret_cast<Handle *>( hh );
        HIP_TRY( hipSetDevice( h.device ) );
        HIP_TRY( hipStreamSynchronize( h.stream ) );
        h.freeSdf();
        h.freeAll();
        h.built     = false;
        h.released  = true;
        h.stageNext = 0;
        return SESDF_OK;
    }

    sesdf_status sesdf_set_neighbor_policy( sesdf_handle * hh, int policy )
    {
        if ( !hh || ( policy != SESDF_NEIGHBORS_TRUNCATE && policy != SESDF_NEIGHBORS_GROW ) )
            return SESDF_ERR_INVALID;
        reinterpret_cast<Handle *>( hh )->growNeighbors = policy == SESDF_NEIGHBORS_GROW;
        return SESDF_OK;
    }

    sesdf_status sesdf_build( sesdf_handle * hh )
    {
        if ( !hh )
            return SESDF_ERR_INVALID;
        Handle & h = *reinterpret_cast<Handle *>( hh );
        HIP_TRY( hipSetDevice( h.device ) );
        BUFFERS_TRY( h );
        if ( h.n )
        {
            if ( !h.hostAtoms )
                return SESDF_ERR_INVALID;
            // #1 CPU => GPU (sesdf.cu:211-213)
            HIP_TRY( hipMemcpyAsync( h.input, h.hostAtoms, size_t( h.n ) * sizeof( float4 ), hipMemcpyHostToDevice, h.stream ) );
        }
        return build_resident( h );
    }

    sesdf_status sesdf_build_device( sesdf_handle * hh, const float * d_xyzr )
    {
        if ( !hh || ( !d_xyzr && reinterpret_cast<Handle *>( hh )->n ) )
            return SESDF_ERR_INVALID;
        Handle & h = *reinterpret_cast<Handle *>( hh );
        HIP_TRY( hipSetDevice( h.device ) );
        BUFFERS_TRY( h );
        if ( h.n && reinterpret_cast<const float4 *>( d_xyzr ) != h.input )
            HIP_TRY( hipMemcpyAsync( h.input, d_xyzr, size_t( h.n ) * sizeof( float4 ), hipMemcpyDeviceToDevice, h.stream ) );
        return build_resident( h );
    }

    sesdf_status sesdf_enqueue( sesdf_handle * hh )
    {
        if ( !hh )
            return SESDF_ERR_INVALID;
        Handle & h = *reinterpret_cast<Handle *>( hh );
        HIP_TRY( hipSetDevice( h.device ) );
        BUFFERS_TRY( h );
        return enqueue( h );
    }

    sesdf_status sesdf_sync( sesdf_handle * hh )
    {
        if ( !hh )
            return SESDF_ERR_INVALID;
        Handle &     h  = *reinterpret_cast<Handle *>( hh );
        sesdf_status st = readback( h );
        if ( st == SESDF_OK )
            h.built = true;
        if ( st == SESDF_OK && overflowed( h ) )
        {
            g_lastError = "capacity overflow in an enqueued build (use sesdf_build)";
            return SESDF_ERR_CAPACITY;
        }
        return st;
    }

    sesdf_status sesdf_build_timed( sesdf_handle * hh, float * stageMs, int stageCount )
    {
        if ( !hh )
            return SESDF_ERR_INVALID;
        Handle & h = *reinterpret_cast<Handle *>( hh );
        HIP_TRY( hipSetDevice( h.device ) );
        BUFFERS_TRY( h );
        for ( auto & e : h.events )
            if ( !e )
                HIP_TRY( hipEventCreate( &e ) );
        if ( h.n == 0 )
            return SESDF_OK;
        h.buildSerial++;
        enqueue_stages( h, h.stream, h.events );
        HIP_TRY( hipGetLastError() );
        sesdf_status st = readback( h );
        if ( st != SESDF_OK )
            return st;
        for ( int q = 0; q < stageCount && q < SESDF_STAGE_COUNT; q++ )
            HIP_TRY( hipEventElapsedTime( stageMs + q, h.events[ q ], h.events[ q + 1 ] ) );
        h.built = true;
        return overflowed( h ) ? SESDF_ERR_CAPACITY : SESDF_OK;
    }

    sesdf_status sesdf_get_counts( sesdf_handle * hh, sesdf_counts * out )
    {
        if ( !hh || !out )
            return SESDF_ERR_INVALID;
        const Handle &  h = *reinterpret_cast<Handle *>( hh );
        const Control & c = *h.hctl;
        std::memset( out, 0, sizeof( *out ) );
        out->atom_nb                      = h.n;
        out->neighbor_nb                  = c.neighborNb;
        out->sector_nb                    = c.sectorNb;
        out->intersected_circle_nb        = c.intersectedCircleNb;
        out->full_circle_nb               = c.fullCircleNb;
        out->visible_circle_nb            = h.mode == 0 ? c.sectorNb : c.visibleCircleNb;
        out->concave_patch_nb             = c.intersectionWithNbNb + c.intersectionNoNbNb;
        out->concave_with_neighbors_nb    = c.intersectionWithNbNb;
        out->segment_nb                   = c.segmentNb;
        out->torus_nb                     = c.torusNb;
        out->neighbor_overflow_atoms      = c.nbOverflowAtoms;
        out->probe_neighbor_overflow      = c.probeNbOverflowNb;
        out->circle_intersection_overflow = c.circleIntOverflowNb;
        out->jk_anomaly_nb                = c.jkAnomalyNb;
        out->fp64_resolve_nb              = c.fp64ResolveNb;
        out->rebuild_nb                   = h.rebuilds;
        out->max_atom_records             = c.maxAtomRecords;
        out->device_bytes                 = h.deviceBytes + h.sdfBytes;
        out->build_serial                 = h.buildSerial;
        out->grid_relayout_nb             = h.relayouts;
        uint32_t compOverflow             = 0;
        if ( h.sdf.compAlloc && h.sdf.compSerial == h.buildSerial &&
             hipMemcpy( &compOverflow, h.sdf.comp.extComp + 1, 4, hipMemcpyDeviceToHost ) != hipSuccess )
            compOverflow = 0;
        out->component_overflow_atoms = compOverflow;
        out->max_neighbors            = h.K;
        out->neighbor_growth_nb       = h.neighborGrowths;
        return SESDF_OK;
    }

    sesdf_status sesdf_get_data( sesdf_handle * hh, sesdf_data * out )
    {
        if ( !hh || !out )
            return SESDF_ERR_INVALID;
        const Handle & h = *reinterpret_cast<Handle *>( hh );
        if ( !h.built )
            return SESDF_ERR_NOT_BUILT;
        const Control & c = *h.hctl;
        std::memset( out, 0, sizeof( *out ) );
        out->atom_nb                   = h.n;
        out->atoms                     = reinterpret_cast<float *>( h.b.atoms );
        out->convex_patch_nb           = h.mode == 0 ? h.n : 0;
        out->convex_patches            = reinterpret_cast<uint32_t *>( h.b.convexPatches );
        out->sector_nb                 = c.sectorNb;
        out->sectors                   = reinterpret_cast<float *>( h.b.sectors );
        out->segment_patch_nb          = c.segmentNb;
        out->segment_patches           = reinterpret_cast<uint32_t *>( h.b.segments );
        out->circle_patch_nb           = c.fullCircleNb;
        out->circle_patches            = reinterpret_cast<uint32_t *>( h.b.fullCircles );
        out->concave_patch_nb          = c.intersectionWithNbNb + c.intersectionNoNbNb;
        out->concave_patches_position  = reinterpret_cast<float *>( h.b.concavePos );
        out->concave_patches_id        = reinterpret_cast<int32_t *>( h.b.concaveIds );
        out->concave_patches_neighbors = reinterpret_cast<float *>( h.b.concaveNb );
        out->probe_radius              = h.rp;
        out->max_concave_neighbors     = uint16_t( h.maxProbeNb );
        return SESDF_OK;
    }

    sesdf_status sesdf_get_exterior_data( sesdf_handle * hh, ssesdf_data * out )
    {
        if ( !hh || !out )
            return SESDF_ERR_INVALID;
        const Handle & h = *reinterpret_cast<Handle *>( hh );
        if ( !h.built )
            return SESDF_ERR_NOT_BUILT;
        const Control & c = *h.hctl;
        std::memset( out, 0, sizeof( *out ) );
        out->atom_nb                   = h.n;
        out->atoms                     = reinterpret_cast<float *>( h.b.atoms );
        out->circle_patch_nb           = c.torusNb;
        out->circle_position           = reinterpret_cast<float *>( h.b.torusPos );
        out->circle_axis               = reinterpret_cast<float *>( h.b.torusAxis );
        out->circle_vs                 = reinterpret_cast<float *>( h.b.torusVs );
        out->concave_patch_nb          = c.intersectionWithNbNb + c.intersectionNoNbNb;
        out->concave_patches_position  = reinterpret_cast<float *>( h.b.concavePos );
        out->concave_patches_id        = reinterpret_cast<int32_t *>( h.b.concaveIds );
        out->concave_patches_neighbors = reinterpret_cast<float *>( h.b.concaveNb );
        out->probe_radius              = h.rp;
        out->max_concave_neighbors     = uint16_t( h.maxProbeNb );
        return SESDF_OK;
    }

    sesdf_status sesdf_get_internal( sesdf_handle * hh, sesdf_internal * out )
    {
        if ( !hh || !out )
            return SESDF_ERR_INVALID;
        const Handle & h = *reinterpret_cast<Handle *>( hh );
        std::memset( out, 0, sizeof( *out ) );
        out->sorted_to_initial            = h.b.sortedToInitial;
        out->neighbor_count               = h.b.nbCount;
        out->neighbor_ids                 = h.b.nbIds;
        out->circle_status                = h.b.status;
        out->trimmed_to_global            = h.b.trimmedToGlobal;
        out->circle_intersection_nb       = h.b.circleIntNb;
        out->max_neighbors                = h.K;
        out->max_intersections_per_circle = 16;
        out->max_intersection_neighbors   = uint16_t( h.maxProbeNb );
        return SESDF_OK;
    }

    sesdf_status sesdf_get_atom_indices( sesdf_handle * hh, uint32_t * hostOut )
    {
        if ( !hh || !hostOut )
            return SESDF_ERR_INVALID;
        Handle & h = *reinterpret_cast<Handle *>( hh );
        if ( !h.built )
            return SESDF_ERR_NOT_BUILT;
        HIP_TRY( hipSetDevice( h.device ) );
        HIP_TRY( hipMemcpyAsync( hostOut, h.b.sortedToInitial, size_t( h.n ) * 4, hipMemcpyDeviceToHost, h.stream ) );
        HIP_TRY( hipStreamSynchronize( h.stream ) );
        return SESDF_OK;
    }

    sesdf_status sesdf_get_graphics_layout( sesdf_handle * hh, sesdf_graphics_layout * out )
    {
        if ( !hh || !out )
            return SESDF_ERR_INVALID;
        const Handle & h = *reinterpret_cast<Handle *>( hh );
        if ( !h.built )
            return SESDF_ERR_NOT_BUILT;
        const Control & c = *h.hctl;
        const uint64_t  N = h.n, I = uint64_t( c.intersectionWithNbNb ) + c.intersectionNoNbNb;
        const uint64_t  nbBytes = uint64_t( c.intersectionWithNbNb ) * h.maxProbeNb * 16;
        std::memset( out, 0, sizeof( *out ) );
        auto span = [ & ]( int q, uint32_t buffer, uint64_t offset, uint64_t size ) { out->spans[ q ] = { buffer, offset, size }; };
        if ( h.mode == 0 )
        {
            const uint64_t Cf = c.fullCircleNb, V = c.sectorNb, G = c.segmentNb;
            const uint64_t sectorOffset = Cf * 8 + ( Cf * 8 ) % 16;
            out->buffer_nb = 5;
            out->buffer_bytes[ 0 ] = N * 16, out->buffer_bytes[ 1 ] = N * 8, out->buffer_bytes[ 2 ] = sectorOffset + V * 16;
            out->buffer_bytes[ 3 ] = G * 16, out->buffer_bytes[ 4 ] = I * 32 + nbBytes;
            span( 0, 0, 0, N * 16 );             // atoms
            span( 1, 1, 0, N * 8 );              // convex patches
            span( 2, 2, sectorOffset, V * 16 );  // sectors
            span( 3, 3, 0, G * 16 );             // segments
            span( 4, 2, 0, Cf * 8 );             // full circles
            span( 5, 4, I * 16, I * 16 );        // concave positions
            span( 6, 4, 0, I * 16 );             // concave ids
            span( 7, 4, I * 32, nbBytes );       // concave neighbours
        }
        else
        {
            const uint64_t T = c.torusNb;
            out->buffer_nb = 3;
            out->buffer_bytes[ 0 ] = N * 16, out->buffer_bytes[ 1 ] = T * 48, out->buffer_bytes[ 2 ] = I * 32 + nbBytes;
            span( 0, 0, 0, N * 16 );
            span( 1, 1, 0, T * 16 );
            span( 2, 1, T * 16, T * 16 );
            span( 3, 1, T * 32, T * 16 );
            span( 4, 2, 0, I * 16 );
            span( 5, 2, I * 16, I * 16 );
            span( 6, 2, I * 32, nbBytes );
        }
        return SESDF_OK;
    }

    sesdf_status sesdf_pack_graphics( sesdf_handle * hh, void * const * d )
    {
        if ( !hh || !d )
            return SESDF_ERR_INVALID;
        Handle &              h = *reinterpret_cast<Handle *>( hh );
        sesdf_graphics_layout L;
        const sesdf_status    st = sesdf_get_graphics_layout( hh, &L );
        if ( st != SESDF_OK )
            return st;
        for ( uint32_t q = 0; q < L.buffer_nb; q++ )
            if ( L.buffer_bytes[ q ] && !d[ q ] )
                return SESDF_ERR_INVALID;
        HIP_TRY( hipSetDevice( h.device ) );
        hipStream_t s   = h.stream;
        auto        put = [ & ]( const sesdf_graphics_span & sp, const void * src ) -> hipError_t {
            if ( !sp.size )
                return hipSuccess;
            return hipMemcpyAsync( static_cast<char *>( d[ sp.buffer ] ) + sp.offset, src, sp.size, hipMemcpyDeviceToDevice, s );
        };
        const Buffers & b = h.b;
        if ( h.mode == 0 )
        {
            const void * src[ 8 ] = { b.atoms, b.convexPatches, b.sectors, b.segments, b.fullCircles, b.concavePos, b.concaveIds, b.concaveNb };
            for ( int q = 0; q < 8; q++ )
                HIP_TRY( put( L.spans[ q ], src[ q ] ) );
            // the alignment gap between full circles and sectors
            const uint64_t gap = L.spans[ 2 ].offset - L.spans[ 4 ].size;
            if ( gap )
                HIP_TRY( hipMemsetAsync( static_cast<char *>( d[ 2 ] ) + L.spans[ 4 ].size, 0, gap, s ) );
        }
        else
        {
            const void * src[ 7 ] = { b.atoms, b.torusPos, b.torusAxis, b.torusVs, b.concavePos, b.concaveIds, b.concaveNb };
            for ( int q = 0; q < 7; q++ )
                HIP_TRY( put( L.spans[ q ], src[ q ] ) );
        }
        HIP_TRY( hipStreamSynchronize( s ) );
        return SESDF_OK;
    }

    sesdf_status sesdf_device_to_host( void * host, const void * device, size_t bytes )
    {
        if ( bytes && ( !host || !device ) )
            return SESDF_ERR_INVALID;
        if ( bytes )
            HIP_TRY( hipMemcpy( host, device, bytes, hipMemcpyDeviceToHost ) );
        return SESDF_OK;
    }

    sesdf_status sesdf_upload( sesdf_handle * hh )
    {
        if ( !hh )
            return SESDF_ERR_INVALID;
        Handle & h = *reinterpret_cast<Handle *>( hh );
        HIP_TRY( hipSetDevice( h.device ) );
        BUFFERS_TRY( h );
        if ( h.n )
        {
            if ( !h.hostAtoms )
                return SESDF_ERR_INVALID;
            HIP_TRY( hipMemcpyAsync( h.input, h.hostAtoms, size_t( h.n ) * sizeof( float4 ), hipMemcpyHostToDevice, h.stream ) );
        }
        return SESDF_OK;
    }

    // Stage entry points: the build's stage sequence in its order, one call per stage of the
    // reference's operations (stream-ordered, no readback; sesdf_sync ends the sequence).
    static sesdf_status run_stage( sesdf_handle * hh, uint32_t stage )
    {
        if ( !hh )
            return SESDF_ERR_INVALID;
        Handle & h = *reinterpret_cast<Handle *>( hh );
        if ( ( stage == 2 && h.mode != 0 ) || ( stage == 4 && h.mode != 1 ) )
            return set_error( SESDF_ERR_UNSUPPORTED, stage == 2 ? "conic patches (segments) belong to the complete surface"
                                                                 : "torus records belong to the exterior path" );
        if ( stage != 0 && stage < h.stageNext )
            return set_error( SESDF_ERR_INVALID, "stages must run in order, starting with sesdf_find_circles" );
        if ( stage != 0 && h.stageNext == 0 )
            return set_error( SESDF_ERR_INVALID, "sesdf_find_circles must run first" );
        HIP_TRY( hipSetDevice( h.device ) );
        BUFFERS_TRY( h );
        const Buffers & b = h.b;
        hipStream_t     s = h.stream;
        if ( h.n )
            switch ( stage )
            {
            case 0:
                launch_begin( b, s );
                launch_grid( b, s );
                launch_circles( b, s );
                launch_circles_finish( b, s );
                break;
            case 1:
                launch_intersections( b, s );
                launch_intersections_finish( b, s );
                launch_circle_scan( b, s );
                launch_fill( b, s );
                break;
            case 2: launch_segments( b, s ); break;
            case 3: launch_probe_neighbors( b, s ); break;
            case 4: launch_torus( b, s ); break;
            }
        if ( stage == 0 )
            h.buildSerial++;
        h.stageNext = stage + 1;
        HIP_TRY( hipGetLastError() );
        return SESDF_OK;
    }
    sesdf_status sesdf_find_circles( sesdf_handle * h ) { return run_stage( h, 0 ); }
    sesdf_status sesdf_find_intersections( sesdf_handle * h ) { return run_stage( h, 1 ); }
    sesdf_status sesdf_build_conic_patches( sesdf_handle * h ) { return run_stage( h, 2 ); }
    sesdf_status sesdf_handle_intersection_singularities( sesdf_handle * h ) { return run_stage( h, 3 ); }
    sesdf_status sesdf_fill_circles( sesdf_handle * h ) { return run_stage( h, 4 ); }

    sesdf_status sesdf_get_global_to_trimmed( sesdf_handle * hh, uint32_t ** d_map, uint64_t * count )
    {
        if ( !hh || !d_map || !count )
            return SESDF_ERR_INVALID;
        Handle & h = *reinterpret_cast<Handle *>( hh );
        if ( !h.built )
            return SESDF_ERR_NOT_BUILT;
        if ( h.mode != 0 )
            return set_error( SESDF_ERR_UNSUPPORTED, "globalToTrimmed belongs to the complete surface" );
        HIP_TRY( hipSetDevice( h.device ) );
        const uint64_t slots = uint64_t( h.n ) * h.K + 1;
        if ( !h.g2t )
        {
            void * p = nullptr;
            HIP_TRY( hipMalloc( &p, slots * 4 ) );
            h.g2t = static_cast<uint32_t *>( p );
        }
        if ( h.g2tSerial != h.buildSerial )
        {
            if ( h.n )
                launch_global_to_trimmed( h.b, h.g2t, h.stream );
            else
                HIP_TRY( hipMemsetAsync( h.g2t, 0, 4, h.stream ) );
            HIP_TRY( hipGetLastError() );
            HIP_TRY( hipStreamSynchronize( h.stream ) );
            h.g2tSerial = h.buildSerial;
        }
        *d_map = h.g2t;
        *count = slots;
        return SESDF_OK;
    }

    sesdf_status sesdf_batch_offsets( sesdf_handle * hh, uint64_t * out )
    {
        if ( !hh || !out )
            return SESDF_ERR_INVALID;
        Handle & h = *reinterpret_cast<Handle *>( hh );
        if ( !h.built )
            return SESDF_ERR_NOT_BUILT;
        HIP_TRY( hipSetDevice( h.device ) );
        const size_t          M = h.mols.size();
        std::vector<uint32_t> bases( M + 1 );
        for ( size_t m = 0; m < M; m++ )
            bases[ m ] = h.mols[ m ].atomBase;
        bases[ M ] = h.n;
        // gather per-atom prefixes at the molecule boundaries
        std::vector<uint2>    cp( h.n );
        std::vector<uint32_t> fb( h.n + 1 ), ws( h.n + 1 ), ns( h.n + 1 ), tb( h.n + 1 ), trim( h.n + 1 );
        std::vector<uint32_t> segStart;
        const Control &       c = *h.hctl;
        if ( h.n )
        {
            HIP_TRY( hipMemcpy( ws.data(), h.b.wStart, ( h.n + 1 ) * 4, hipMemcpyDeviceToHost ) );
            HIP_TRY( hipMemcpy( ns.data(), h.b.nStart, ( h.n + 1 ) * 4, hipMemcpyDeviceToHost ) );
            if ( h.mode == 0 )
            {
                HIP_TRY( hipMemcpy( cp.data(), h.b.convexPatches, h.n * 8, hipMemcpyDeviceToHost ) );
                HIP_TRY( hipMemcpy( fb.data(), h.b.fullBase, h.n * 4, hipMemcpyDeviceToHost ) );
                HIP_TRY( hipMemcpy( trim.data(), h.b.trimBase, h.n * 4, hipMemcpyDeviceToHost ) );
                segStart.resize( size_t( std::min( c.intersectedCircleNb, h.cap.trimmed ) ) + 1 );
                HIP_TRY( hipMemcpy( segStart.data(), h.b.segStart, segStart.size() * 4, hipMemcpyDeviceToHost ) );
            }
            else
                HIP_TRY( hipMemcpy( tb.data(), h.b.torusBase, ( h.n + 1 ) * 4, hipMemcpyDeviceToHost ) );
        }
        for ( size_t m = 0; m <= M; m++ )
        {
            const uint32_t a   = bases[ m ];
            const bool     end = a >= h.n;
            uint64_t *     o   = out;
            o[ 0 * ( M + 1 ) + m ] = a;
            o[ 1 * ( M + 1 ) + m ] = h.mode == 0 ? ( end ? c.sectorNb : cp[ a ].x ) : 0;
            o[ 2 * ( M + 1 ) + m ] = h.mode == 0 ? ( end ? c.fullCircleNb : fb[ a ] ) : 0;
            o[ 3 * ( M + 1 ) + m ] = end ? c.intersectionWithNbNb : ws[ a ];
            o[ 4 * ( M + 1 ) + m ] = end ? c.intersectionNoNbNb : ns[ a ];
            if ( h.mode == 0 )
            {
                const uint32_t t = end ? uint32_t( segStart.size() - 1 ) : std::min<uint32_t>( trim[ a ], uint32_t( segStart.size() - 1 ) );
                o[ 5 * ( M + 1 ) + m ] = segStart.empty() ? 0 : segStart[ t ];
            }
            else
                o[ 5 * ( M + 1 ) + m ] = 0;
            o[ 6 * ( M + 1 ) + m ] = h.mode == 1 ? ( end ? c.torusNb : tb[ a ] ) : 0;
        }
        return SESDF_OK;
    }
}
