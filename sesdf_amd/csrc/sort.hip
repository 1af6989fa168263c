// Stable LSD radix sort of (key, value) pairs: per 8-bit digit, a histogram per tile of 2048 items,
// a scan of each digit's counts over the tiles, and a scatter that keeps the order of equal digits.
// The scatter ranks the items of a wave by their digit with eight ballots (lanes of equal digit
// form one mask), so no LDS atomics order the writes and the result is deterministic.
#include "sort.h"

#include "common.h"

#include <utility>

namespace sesdf
{
    namespace
    {
        constexpr int Threads = 256;
        constexpr int Items   = int( SortTile ) / Threads;
        constexpr int Waves   = Threads / WaveSize;

        __global__ void __launch_bounds__( Threads ) sort_hist_kernel( const uint32_t * keys, uint32_t n, int shift, uint32_t * tileHist,
                                                                      uint32_t tiles )
        {
            __shared__ uint32_t h[ 256 ];
            h[ threadIdx.x ] = 0;
            __syncthreads();
            const uint64_t t0 = uint64_t( blockIdx.x ) * SortTile;
            for ( int r = 0; r < Items; r++ )
            {
                const uint64_t q = t0 + uint64_t( r ) * Threads + threadIdx.x;
                if ( q < n )
                    atomicAdd( &h[ ( keys[ q ] >> shift ) & 255u ], 1u );
            }
            __syncthreads();
            tileHist[ size_t( threadIdx.x ) * tiles + blockIdx.x ] = h[ threadIdx.x ];
        }

        // Exclusive scan of one digit's tile counts (one workgroup per digit); the digit's total goes
        // to digitTotal.
        __global__ void __launch_bounds__( Threads ) sort_scan_kernel( uint32_t * tileHist, uint32_t tiles, uint32_t * digitTotal )
        {
            __shared__ uint32_t wsum[ Waves ];
            uint32_t *          row   = tileHist + size_t( blockIdx.x ) * tiles;
            const int           w     = int( threadIdx.x ) / WaveSize;
            uint32_t            carry = 0;
            for ( uint32_t base = 0; base < tiles; base += Threads )
            {
                const uint32_t q    = base + threadIdx.x;
                const uint32_t v    = q < tiles ? row[ q ] : 0u;
                const uint32_t incl = wave_incl_scan( v );
                if ( lane_id() == WaveSize - 1 )
                    wsum[ w ] = incl;
                __syncthreads();
                uint32_t pre = 0, all = 0;
                for ( int x = 0; x < Waves; x++ )
                {
                    pre += x < w ? wsum[ x ] : 0u;
                    all += wsum[ x ];
                }
                if ( q < tiles )
                    row[ q ] = carry + pre + incl - v;
                carry += all;
                __syncthreads();
            }
            if ( threadIdx.x == 0 )
                digitTotal[ blockIdx.x ] = carry;
        }

        __global__ void __launch_bounds__( Threads ) sort_scatter_kernel( const uint32_t * kin, const uint32_t * vin, uint32_t * kout,
                                                                         uint32_t * vout, uint32_t n, int shift, const uint32_t * tileHist,
                                                                         uint32_t tiles, const uint32_t * digitTotal )
        {
            __shared__ uint32_t base[ 256 ];
            __shared__ uint32_t wcnt[ Waves ][ 256 ];
            __shared__ uint32_t wsum[ Waves ];
            const int           tid  = int( threadIdx.x );
            const int           w    = tid / WaveSize;
            const int           lane = lane_id();
            // first output slot of each digit in this tile: the digits before it, then the tiles before
            const uint32_t tot  = digitTotal[ tid ];
            const uint32_t incl = wave_incl_scan( tot );
            if ( lane == WaveSize - 1 )
                wsum[ w ] = incl;
            __syncthreads();
            uint32_t pre = 0;
            for ( int x = 0; x < w; x++ )
                pre += wsum[ x ];
            base[ tid ] = pre + incl - tot + tileHist[ size_t( tid ) * tiles + blockIdx.x ];
            const uint64_t t0 = uint64_t( blockIdx.x ) * SortTile;
            for ( int r = 0; r < Items; r++ )
            {
                for ( int x = 0; x < Waves; x++ )
                    wcnt[ x ][ tid ] = 0;
                __syncthreads();
                const uint64_t q     = t0 + uint64_t( r ) * Threads + tid;
                const bool     valid = q < n;
                const uint32_t key   = valid ? kin[ q ] : 0u;
                const uint32_t d     = ( key >> shift ) & 255u;
                // lanes holding the same digit: eight ballots
                uint64_t match = ballot( valid );
                for ( int bit = 0; bit < 8; bit++ )
                {
                    const bool     set = ( d >> bit ) & 1u;
                    const uint64_t bb  = ballot( set );
                    match &= set ? bb : ~bb;
                }
                const uint32_t rank = rank_below( match );
                if ( valid && ( match >> lane ) == 1u )
                    wcnt[ w ][ d ] = uint32_t( __popcll( match ) ); // the group's last lane
                __syncthreads();
                if ( valid )
                {
                    uint32_t off = base[ d ] + rank;
                    for ( int x = 0; x < w; x++ )
                        off += wcnt[ x ][ d ];
                    kout[ off ] = key;
                    vout[ off ] = vin[ q ];
                }
                __syncthreads();
                uint32_t add = 0;
                for ( int x = 0; x < Waves; x++ )
                    add += wcnt[ x ][ tid ];
                base[ tid ] += add;
            }
        }
    } // namespace

    int radix_sort_pairs( uint32_t * k0, uint32_t * v0, uint32_t * k1, uint32_t * v1, uint32_t * hist, uint32_t n, int bits,
                          hipStream_t stream )
    {
        if ( n == 0 )
            return 0;
        const uint32_t tiles      = uint32_t( ( uint64_t( n ) + SortTile - 1 ) / SortTile );
        uint32_t *     digitTotal = hist + size_t( 256 ) * tiles;
        uint32_t *     kin = k0, *vin = v0, *kout = k1, *vout = v1;
        int            where = 0;
        for ( int shift = 0; shift < bits; shift += 8 )
        {
            sort_hist_kernel<<<tiles, Threads, 0, stream>>>( kin, n, shift, hist, tiles );
            sort_scan_kernel<<<256, Threads, 0, stream>>>( hist, tiles, digitTotal );
            sort_scatter_kernel<<<tiles, Threads, 0, stream>>>( kin, vin, kout, vout, n, shift, hist, tiles, digitTotal );
            std::swap( kin, kout );
            std::swap( vin, vout );
            where ^= 1;
        }
        return where;
    }
} // namespace sesdf
