// Uniform-grid build: counting sort of atoms by cell (stable in input order), one pass per step.
//
// Replaces AccelerationGrid::build (sesdf/src/cuda/grid.cu:99-132: hash -> radix sort -> cell
// start/end -> gather). Here the cell table is a dense per-molecule brick whose edge is chosen by
// the caller (not the reference's fixed 16^3 / 32^3 heuristic): histogram with atomics, decoupled
// look-back scan of the counts, scatter, then an in-cell rank by input index so that the sorted
// order (and therefore every output order) is deterministic.
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
        const uint32_t a     = blockIdx.x * blockDim.x + threadIdx.x;
        float          r     = 0.f;
        if ( a < b.n )
        {
            const float4   p = b.input[ a ];
            const uint32_t m = find_molecule( b.mols, b.molNb, a );
            const Molecule M = b.mols[ m ];
            int            gx, gy, gz;
            cell_of( M, p.x, p.y, p.z, gx, gy, gz );
            const uint32_t c = cell_index( M, gx, gy, gz );
            b.atomCell[ a ]  = c;
            atomicAdd( b.cellCount + c, 1u );
            r = fmaxf( p.w, 0.f );
        }
        // wave max of the radii, one atomic per wave
        for ( int o = 32; o > 0; o >>= 1 )
            r = fmaxf( r, __shfl_xor( r, o, WaveSize ) );
        if ( lane_id() == 0 )
            atomicMax( &b.ctl->radiusMaxBits, __float_as_uint( r ) );
    }

    // Exclusive scan of cellCount into cellStart (tiles of 2048 cells, ticketed look-back);
    // clears cellCount for the next build.
    constexpr int ScanThreads = 256;
    constexpr int ScanItems   = 8;
    constexpr int ScanTile    = ScanThreads * ScanItems;

    __global__ void __launch_bounds__( ScanThreads ) cell_scan_kernel( Buffers b )
    {
        __shared__ uint32_t waveSums[ ScanThreads / WaveSize ];
        __shared__ uint32_t tileBase;
        __shared__ uint32_t ticket;
        const uint32_t      tiles = ( b.cellNb + ScanTile - 1 ) / ScanTile;
        const uint32_t      tag   = stage_tag( b.ctl, 1 );
        for ( ;; )
        {
            if ( threadIdx.x == 0 )
                ticket = atomicAdd( &b.ctl->tickets[ TicketCellScan ], 1u );
            __syncthreads();
            const uint32_t t = ticket;
            if ( t >= tiles )
                break;
            const uint32_t base = t * ScanTile + threadIdx.x * ScanItems;
            uint32_t       v[ ScanItems ];
            uint32_t       sum = 0;
#pragma unroll
            for ( int q = 0; q < ScanItems; q++ )
            {
                v[ q ] = base + q < b.cellNb ? b.cellCount[ base + q ] : 0u;
                sum += v[ q ];
            }
            const uint32_t incl = wave_incl_scan( sum );
            const int      w    = threadIdx.x / WaveSize;
            if ( lane_id() == WaveSize - 1 )
                waveSums[ w ] = incl;
            __syncthreads();
            if ( w == 0 )
            {
                uint32_t tot = 0;
                for ( int q = 0; q < ScanThreads / WaveSize; q++ )
                    tot += waveSums[ q ];
                uint32_t local[ 1 ] = { tot }, excl[ 1 ];
                wave_lookback<1>( b.flags, b.flagStride, t, tag, local, excl );
                if ( lane_id() == 0 )
                {
                    tileBase = excl[ 0 ];
                    if ( t == tiles - 1 )
                        b.cellStart[ b.cellNb ] = excl[ 0 ] + tot;
                }
            }
            __syncthreads();
            uint32_t run = tileBase + incl - sum;
            for ( int q = 0; q < w; q++ )
                run += waveSums[ q ];
#pragma unroll
            for ( int q = 0; q < ScanItems; q++ )
                if ( base + q < b.cellNb )
                {
                    b.cellStart[ base + q ] = run;
                    b.cellCount[ base + q ] = 0u;
                    run += v[ q ];
                }
            __syncthreads();
        }
    }

    __global__ void __launch_bounds__( 256 ) grid_scatter_kernel( Buffers b )
    {
        const uint32_t a = blockIdx.x * blockDim.x + threadIdx.x;
        if ( a >= b.n )
            return;
        const uint32_t c    = b.atomCell[ a ];
        const uint32_t slot = b.cellStart[ c ] + atomicAdd( b.cellFill + c, 1u );
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
        const uint32_t dst      = st + rank;
        b.sortedToInitial[ dst ] = id;
        b.atoms[ dst ]           = b.input[ id ];
        if ( s == st )
            b.cellFill[ c ] = 0u;
    }

    static inline uint32_t blocks_for( uint64_t n, uint32_t threads ) { return uint32_t( ( n + threads - 1 ) / threads ); }

    void launch_begin( const Buffers & b, hipStream_t s ) { begin_kernel<<<1, 64, 0, s>>>( b.ctl ); }

    void launch_grid( const Buffers & b, hipStream_t s )
    {
        if ( b.n == 0 )
            return;
        grid_hash_kernel<<<blocks_for( b.n, 256 ), 256, 0, s>>>( b );
        const uint32_t tiles = ( b.cellNb + ScanTile - 1 ) / ScanTile;
        const uint32_t grid  = tiles < 1024u ? tiles : 1024u;
        cell_scan_kernel<<<grid, ScanThreads, 0, s>>>( b );
        grid_scatter_kernel<<<blocks_for( b.n, 256 ), 256, 0, s>>>( b );
        grid_rank_kernel<<<blocks_for( b.n, 256 ), 256, 0, s>>>( b );
    }
} // namespace sesdf
