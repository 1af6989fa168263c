// Uniform-grid build (counting sort of atoms by cell, stable in input order) and the generic
// tile scan used by every stage for its per-item offsets.
//
// Replaces AccelerationGrid::build (sesdf/src/cuda/grid.cu:99-132: hash -> radix sort -> cell
// start/end -> gather) and the thrust scans / copy_if / count_if of the reference stages
// (sesdf/include/bcs/sesdf/operations.inl:46-70, 149-156, 500-517; ssesdf/src/ssesdf/operations.cu:44-54).
// The cell table is a dense per-molecule brick whose edge is chosen by the caller (not the
// reference's fixed 16^3 / 32^3 heuristic): histogram with atomics, decoupled look-back scan of
// the counts, scatter, then an in-cell rank by input index so that the sorted order (and every
// output order) is deterministic.
#include "kernels.h"

namespace sesdf
{
    __global__ void __launch_bounds__( 64 ) begin_kernel( Control * ctl )
    {
        if ( threadIdx.x == 0 )
        {
            const uint32_t epoch = ctl->epoch + 1;
            uint32_t *     words = reinterpret_cast<uint32_t *>( ctl );
            for ( uint32_t w = 1; w < sizeof( Control ) / 4; w++ )
                words[ w ] = 0;
            ctl->epoch = epoch;
        }
    }

    __global__ void __launch_bounds__( 256 ) grid_hash_kernel( Buffers b )
    {
        __shared__ float waveMax[ 4 ];
        const uint32_t   a = blockIdx.x * blockDim.x + threadIdx.x;
        float            r = 0.f;
        if ( a < b.n )
        {
            const float4   p = b.input[ a ];
            const uint32_t m = find_molecule( b.mols, b.molNb, a );
            const Molecule M = b.mols[ m ];
            if ( b.molNb > 1 )
                b.atomMol[ a ] = m; // molecule-major order: the same for the sorted index a
            // an atom outside the brick (atoms moved since the grid was laid out, or not finite)
            // would be clamped into a border cell, where the stencil bounds no longer hold: the
            // host lays the grid out again and rebuilds
            const float fx = ( p.x - M.ox ) / M.cell, fy = ( p.y - M.oy ) / M.cell, fz = ( p.z - M.oz ) / M.cell;
            if ( !( fx >= 0.f && fx < float( M.dx ) && fy >= 0.f && fy < float( M.dy ) && fz >= 0.f && fz < float( M.dz ) ) )
                set_overflow( b.ctl, OverflowGrid );
            int            gx, gy, gz;
            cell_of( M, p.x, p.y, p.z, gx, gy, gz );
            const uint32_t c = cell_index( M, gx, gy, gz );
            b.atomCell[ a ]  = c;
            // the arrival rank in the cell places the atom in the scatter (no second atomic pass);
            // atomCnt is free until the circle stage
            b.atomCnt[ a ] = atomicAdd( b.cellCount + c, 1u );
            r = fmaxf( p.w, 0.f );
        }
        for ( int o = 32; o > 0; o >>= 1 )
            r = fmaxf( r, __shfl_xor( r, o, WaveSize ) );
        if ( lane_id() == 0 )
            waveMax[ threadIdx.x / WaveSize ] = r;
        __syncthreads();
        if ( threadIdx.x == 0 )
        {
            const float m = fmaxf( fmaxf( waveMax[ 0 ], waveMax[ 1 ] ), fmaxf( waveMax[ 2 ], waveMax[ 3 ] ) );
            atomicMax( &b.ctl->radiusMaxBits, __float_as_uint( m ) ); // one atomic per block
        }
    }

    constexpr int ScanThreads = 256;

    // Exclusive scan, one tile of ScanThreads x ITEMS items per ticket (tiles in ticket order, so
    // a tile only waits for tiles already being worked on). Each thread owns ITEMS consecutive
    // items. TR: the tile's columns are read and written coalesced (item q * 256 + thread) and
    // transposed through LDS (padded: one word per 32); without it each thread reads its own run.
    // Large arrays take large tiles (fewer look-back steps and tickets).
    template<int NC, int ITEMS, bool TR>
    __global__ void __launch_bounds__( ScanThreads ) scan_kernel( ScanArgs a, Control * ctl, uint64_t * flags, uint32_t flagStride )
    {
        constexpr int       T = ScanThreads, W = T / WaveSize, TILE = T * ITEMS;
        __shared__ uint32_t lds[ TR ? TILE + TILE / 32 : 1 ];
        __shared__ uint32_t waveSums[ NC ][ W ];
        __shared__ uint32_t tileBase[ NC ];
        __shared__ uint32_t ticket;
        __shared__ uint32_t redMax[ W ], redSum[ W ];
        const uint32_t      n     = a.countPtr ? min( *a.countPtr, a.n ) : a.n;
        const uint32_t      tiles = ( n + TILE - 1 ) / TILE;
        const uint32_t      tag   = stage_tag( ctl, a.stage );
        const int           w     = threadIdx.x / WaveSize;
        const int           tid   = threadIdx.x;
        auto                pad   = []( int i ) { return i + ( i >> 5 ); };
        if ( n == 0 )
        {
            if ( blockIdx.x == 0 && threadIdx.x == 0 )
                for ( int c = 0; c < NC; c++ )
                {
                    a.out[ c ][ 0 ] = 0;
                    if ( a.totals[ c ] )
                        *a.totals[ c ] = 0;
                }
            return;
        }
        for ( ;; )
        {
            if ( threadIdx.x == 0 )
                ticket = atomicAdd( &ctl->tickets[ a.ticket ], 1u );
            __syncthreads();
            const uint32_t t = ticket;
            if ( t >= tiles )
                break;
            const uint32_t tile0 = t * TILE;
            const uint32_t base  = tile0 + tid * ITEMS;
            uint32_t       v[ NC ][ ITEMS ];
            uint32_t       s[ NC ], incl[ NC ];
            uint32_t       mx = 0, sm = 0;
            if constexpr ( TR )
            {
#pragma unroll
                for ( int c = 0; c < NC; c++ )
                {
                    s[ c ] = 0;
                    if ( a.halfSecond && c == 1 ) // halves of column 0, already in registers
                    {
#pragma unroll
                        for ( int q = 0; q < ITEMS; q++ )
                        {
                            v[ c ][ q ] = v[ 0 ][ q ] / 2;
                            s[ c ] += v[ c ][ q ];
                        }
                        continue;
                    }
#pragma unroll
                    for ( int q = 0; q < ITEMS; q++ )
                    {
                        const uint32_t i          = tile0 + q * T + tid;
                        lds[ pad( q * T + tid ) ] = i < n ? a.in[ c ][ i ] : 0u;
                    }
                    __syncthreads();
#pragma unroll
                    for ( int q = 0; q < ITEMS; q++ )
                    {
                        v[ c ][ q ] = lds[ pad( tid * ITEMS + q ) ];
                        s[ c ] += v[ c ][ q ];
                    }
                    __syncthreads();
                }
                if ( a.maxOut )
#pragma unroll
                    for ( int q = 0; q < ITEMS; q++ )
                        mx = max( mx, v[ 0 ][ q ] + ( NC > 1 && !a.halfSecond ? v[ 1 ][ q ] : 0u ) );
                if ( a.sumOut ) // a plain sum: read in coalesced order
#pragma unroll
                    for ( int q = 0; q < ITEMS; q++ )
                    {
                        const uint32_t i = tile0 + q * T + tid;
                        sm += i < n ? a.in[ NC ][ i ] : 0u;
                    }
            }
            else
            {
#pragma unroll
                for ( int c = 0; c < NC; c++ )
                    s[ c ] = 0;
#pragma unroll
                for ( int q = 0; q < ITEMS; q++ )
                {
                    const bool in = base + q < n;
#pragma unroll
                    for ( int c = 0; c < NC; c++ )
                    {
                        uint32_t x = 0;
                        if ( in )
                            x = ( a.halfSecond && c == 1 ) ? v[ 0 ][ q ] / 2 : a.in[ c ][ base + q ];
                        v[ c ][ q ] = x;
                        s[ c ] += x;
                    }
                    if ( in && a.maxOut )
                        mx = max( mx, v[ 0 ][ q ] + ( NC > 1 && !a.halfSecond ? v[ 1 ][ q ] : 0u ) );
                    if ( in && a.sumOut )
                        sm += a.in[ NC ][ base + q ];
                }
            }
#pragma unroll
            for ( int c = 0; c < NC; c++ )
            {
                incl[ c ] = wave_incl_scan( s[ c ] );
                if ( lane_id() == WaveSize - 1 )
                    waveSums[ c ][ w ] = incl[ c ];
            }
            if ( a.maxOut || a.sumOut )
            {
                for ( int o = 32; o > 0; o >>= 1 )
                {
                    mx = max( mx, __shfl_xor( mx, o, WaveSize ) );
                    sm += __shfl_xor( sm, o, WaveSize );
                }
                if ( lane_id() == 0 )
                    redMax[ w ] = mx, redSum[ w ] = sm;
            }
            __syncthreads();
            if ( w == 0 )
            {
                uint32_t local[ NC ], excl[ NC ];
#pragma unroll
                for ( int c = 0; c < NC; c++ )
                {
                    local[ c ] = 0;
                    for ( int q = 0; q < W; q++ )
                        local[ c ] += waveSums[ c ][ q ];
                }
                if constexpr ( TR )
                    wave_lookback_joint<NC>( flags, flagStride, t, tag, local, excl );
                else
                    wave_lookback<NC>( flags, flagStride, t, tag, local, excl );
                if ( lane_id() == 0 )
                {
#pragma unroll
                    for ( int c = 0; c < NC; c++ )
                    {
                        tileBase[ c ] = excl[ c ];
                        if ( t == tiles - 1 )
                        {
                            a.out[ c ][ n ] = excl[ c ] + local[ c ];
                            if ( a.totals[ c ] )
                                *a.totals[ c ] = excl[ c ] + local[ c ];
                        }
                    }
                    uint32_t m = 0, total = 0;
                    for ( int q = 0; q < W; q++ )
                        m = max( m, redMax[ q ] ), total += redSum[ q ];
                    if ( a.maxOut )
                        atomicMax( a.maxOut, m );
                    if ( a.sumOut )
                        atomicAdd( a.sumOut, total );
                }
            }
            __syncthreads();
#pragma unroll
            for ( int c = 0; c < NC; c++ )
            {
                uint32_t run = tileBase[ c ] + incl[ c ] - s[ c ];
                for ( int q = 0; q < w; q++ )
                    run += waveSums[ c ][ q ];
                if constexpr ( TR )
                {
#pragma unroll
                    for ( int q = 0; q < ITEMS; q++ )
                    {
                        lds[ pad( tid * ITEMS + q ) ] = run;
                        run += v[ c ][ q ];
                    }
                    __syncthreads();
#pragma unroll
                    for ( int q = 0; q < ITEMS; q++ )
                    {
                        const uint32_t i = tile0 + q * T + tid;
                        if ( i < n )
                            a.out[ c ][ i ] = lds[ pad( q * T + tid ) ];
                    }
                    __syncthreads();
                }
                else
                {
#pragma unroll
                    for ( int q = 0; q < ITEMS; q++ )
                        if ( base + q < n )
                        {
                            a.out[ c ][ base + q ] = run;
                            run += v[ c ][ q ];
                        }
                }
            }
            if ( a.clearInput || a.clearAlso )
#pragma unroll
                for ( int q = 0; q < ITEMS; q++ )
                {
                    const uint32_t i = tile0 + q * T + tid;
                    if ( i < n )
                    {
                        if ( a.clearInput )
                            for ( int c = 0; c < NC; c++ )
                                if ( !( a.halfSecond && c == 1 ) )
                                    a.in[ c ][ i ] = 0u;
                        if ( a.clearAlso )
                            a.clearAlso[ i ] = 0u;
                    }
                }
            __syncthreads();
        }
    }

    template<int NC, int ITEMS, bool TR>
    static void launch_scan_tiles( const Buffers & b, const ScanArgs & a, hipStream_t s )
    {
        const uint32_t tile  = uint32_t( ScanThreads * ITEMS );
        const uint32_t tiles = ( a.n + tile - 1 ) / tile;
        const uint32_t cap   = uint32_t( device_cu_count() ) * 4u;
        const uint32_t grid  = tiles < 1 ? 1 : ( tiles < cap ? tiles : cap );
        scan_kernel<NC, ITEMS, TR><<<grid, ScanThreads, 0, s>>>( a, b.ctl, b.flags, b.flagStride );
    }

    template<int NC>
    void launch_scan( const Buffers & b, const ScanArgs & a, hipStream_t s )
    {
        // a.n bounds the item count (a device-side count may lower it): tiles of 2048 up to 768k
        // items, 4096 (coalesced through LDS) beyond
        static_assert( ScanThreads * 8 == ScanMinTile, "flag arrays are sized for the smallest tile" );
        if ( a.n <= 768u * 1024u )
            launch_scan_tiles<NC, 8, false>( b, a, s );
        else
            launch_scan_tiles<NC, 16, true>( b, a, s );
    }
    template void launch_scan<1>( const Buffers &, const ScanArgs &, hipStream_t );
    template void launch_scan<2>( const Buffers &, const ScanArgs &, hipStream_t );
    template void launch_scan<3>( const Buffers &, const ScanArgs &, hipStream_t );

    __global__ void __launch_bounds__( 256 ) grid_scatter_kernel( Buffers b )
    {
        const uint32_t a = blockIdx.x * blockDim.x + threadIdx.x;
        if ( a >= b.n )
            return;
        const uint32_t c    = b.atomCell[ a ];
        const uint32_t slot = b.cellStart[ c ] + b.atomCnt[ a ];
        b.slotInput[ slot ] = a;
    }

    // Rank every atom of a cell by input index (stable order), gather the sorted atoms.
    __global__ void __launch_bounds__( 256 ) grid_rank_kernel( Buffers b )
    {
        const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
        if ( s >= b.n )
            return;
        const uint32_t id   = b.slotInput[ s ];
        const uint32_t c    = b.atomCell[ id ];
        const uint32_t st   = b.cellStart[ c ];
        const uint32_t en   = b.cellStart[ c + 1 ];
        uint32_t       rank = 0;
        for ( uint32_t q = st; q < en; q++ )
            rank += b.slotInput[ q ] < id;
        const uint32_t dst       = st + rank;
        b.sortedToInitial[ dst ] = id;
        b.atoms[ dst ]           = b.input[ id ];
    }

    static inline uint32_t blocks_for( uint64_t n, uint32_t threads ) { return uint32_t( ( n + threads - 1 ) / threads ); }

    void launch_begin( const Buffers & b, hipStream_t s ) { begin_kernel<<<1, 64, 0, s>>>( b.ctl ); }

    void launch_grid( const Buffers & b, hipStream_t s )
    {
        if ( b.n == 0 )
            return;
        grid_hash_kernel<<<blocks_for( b.n, 256 ), 256, 0, s>>>( b );
        ScanArgs a {};
        a.in[ 0 ]    = b.cellCount;
        a.out[ 0 ]   = b.cellStart;
        a.n          = b.cellNb;
        a.ticket     = TicketCellScan;
        a.stage      = 1;
        a.clearInput = 1;
        launch_scan<1>( b, a, s );
        grid_scatter_kernel<<<blocks_for( b.n, 256 ), 256, 0, s>>>( b );
        grid_rank_kernel<<<blocks_for( b.n, 256 ), 256, 0, s>>>( b );
    }
} // namespace sesdf
