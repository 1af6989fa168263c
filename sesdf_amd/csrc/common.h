// Device-side building blocks shared by every kernel of the SES builder (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sesdf
{
    constexpr int      WaveSize     = 64;
    constexpr uint32_t EmptyCell    = 0xffffffffu;
    constexpr float    PredicateEps = 1e-4f;       // slack of every threshold test of the reference
    constexpr float    TwoPi        = 6.283185307f; // bcs::TwoPif (sesdf/include/bcs/core/math.hpp)

    // ------------------------------------------------------------------------------------------
    // Small vector math (written for this library; FP32 and FP64 variants)
    // ------------------------------------------------------------------------------------------
    struct f3
    {
        float x, y, z;
    };
    struct d3
    {
        double x, y, z;
    };
    __device__ __forceinline__ f3    mk( float x, float y, float z ) { return { x, y, z }; }
    __device__ __forceinline__ f3    xyz( float4 a ) { return { a.x, a.y, a.z }; }
    __device__ __forceinline__ f3    operator+( f3 a, f3 b ) { return { a.x + b.x, a.y + b.y, a.z + b.z }; }
    __device__ __forceinline__ f3    operator-( f3 a, f3 b ) { return { a.x - b.x, a.y - b.y, a.z - b.z }; }
    __device__ __forceinline__ f3    operator-( f3 a ) { return { -a.x, -a.y, -a.z }; }
    __device__ __forceinline__ f3    operator*( f3 a, float s ) { return { a.x * s, a.y * s, a.z * s }; }
    __device__ __forceinline__ f3    operator*( float s, f3 a ) { return { a.x * s, a.y * s, a.z * s }; }
    __device__ __forceinline__ float dot( f3 a, f3 b ) { return a.x * b.x + a.y * b.y + a.z * b.z; }
    __device__ __forceinline__ float len2( f3 a ) { return dot( a, a ); }
    __device__ __forceinline__ f3    cross( f3 a, f3 b ) { return { a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x }; }
    __device__ __forceinline__ f3    unit( f3 v ) { return v * ( 1.0f / sqrtf( dot( v, v ) ) ); }

    // Hardware square root, reciprocal and reciprocal square root (about 1 ulp), where a result
    // only feeds comparisons far wider than that: the correctly rounded forms expand to over ten
    // instructions each.
    __device__ __forceinline__ float fsqrt( float x ) { return __builtin_amdgcn_sqrtf( x ); }
    __device__ __forceinline__ float frcp( float x ) { return __builtin_amdgcn_rcpf( x ); }
    __device__ __forceinline__ float frsq( float x ) { return __builtin_amdgcn_rsqf( x ); }

    __device__ __forceinline__ d3     mkd( f3 a ) { return { a.x, a.y, a.z }; }
    __device__ __forceinline__ d3     operator+( d3 a, d3 b ) { return { a.x + b.x, a.y + b.y, a.z + b.z }; }
    __device__ __forceinline__ d3     operator-( d3 a, d3 b ) { return { a.x - b.x, a.y - b.y, a.z - b.z }; }
    __device__ __forceinline__ d3     operator*( d3 a, double s ) { return { a.x * s, a.y * s, a.z * s }; }
    __device__ __forceinline__ double dot( d3 a, d3 b ) { return a.x * b.x + a.y * b.y + a.z * b.z; }
    __device__ __forceinline__ d3     cross( d3 a, d3 b ) { return { a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x }; }
    __device__ __forceinline__ d3     unit( d3 v ) { return v * ( 1.0 / sqrt( dot( v, v ) ) ); }

    // ------------------------------------------------------------------------------------------
    // Wave64 primitives
    // ------------------------------------------------------------------------------------------
    __device__ __forceinline__ int      lane_id() { return __lane_id(); }
    __device__ __forceinline__ uint64_t ballot( bool p ) { return __builtin_amdgcn_ballot_w64( p ); }
    __device__ __forceinline__ uint64_t lanes_below() { return ( uint64_t( 1 ) << lane_id() ) - 1; }
    __device__ __forceinline__ uint32_t rank_below( uint64_t mask ) { return uint32_t( __popcll( mask & lanes_below() ) ); }

    // Value of lane `src` (uniform) in every lane, as a scalar: the compiler then knows it is
    // wave-uniform, so loops bounded by it stay uniform (no exec-mask bookkeeping).
    template<class T>
    __device__ __forceinline__ T wave_bcast( T v, int src )
    {
        static_assert( sizeof( T ) == 4, "32-bit values" );
        return __builtin_bit_cast( T, __builtin_amdgcn_readlane( __builtin_bit_cast( int, v ), src ) );
    }
    // Wave index of the calling lane within its workgroup, as a uniform value.
    __device__ __forceinline__ int wave_in_block() { return __builtin_amdgcn_readfirstlane( int( threadIdx.x ) / WaveSize ); }
    __device__ __forceinline__ uint32_t wave_sum( uint32_t v )
    {
        for ( int o = 32; o > 0; o >>= 1 )
            v += __shfl_xor( v, o, WaveSize );
        return v;
    }
    __device__ __forceinline__ uint32_t wave_incl_scan( uint32_t v )
    {
        const int l = lane_id();
        for ( int o = 1; o < WaveSize; o <<= 1 )
        {
            const uint32_t t = __shfl_up( v, o, WaveSize );
            if ( l >= o )
                v += t;
        }
        return v;
    }
    // Inclusive prefix maximum over the wave by DPP (row shifts by 1, 2, 4, 8, then the row 15 and
    // lane 31 broadcasts), inside the VALU. A lane without a DPP source keeps its own value, and
    // a maximum may take in any earlier lane more than once, so full row masks are exact. Every
    // lane must be active.
    template<int Ctrl>
    __device__ __forceinline__ int dpp_from( int v )
    {
        return __builtin_amdgcn_update_dpp( v, v, Ctrl, 0xf, 0xf, false );
    }
    __device__ __forceinline__ int wave_incl_max( int v )
    {
        v = max( v, dpp_from<0x111>( v ) ); // row_shr:1
        v = max( v, dpp_from<0x112>( v ) ); // row_shr:2
        v = max( v, dpp_from<0x114>( v ) ); // row_shr:4
        v = max( v, dpp_from<0x118>( v ) ); // row_shr:8
        v = max( v, dpp_from<0x142>( v ) ); // row_bcast:15
        v = max( v, dpp_from<0x143>( v ) ); // row_bcast:31
        return v;
    }

    // ------------------------------------------------------------------------------------------
    // Device control block: tickets, counts and overflow flags of one build. Reset by the first
    // kernel of every build; read back by the host once at the end.
    // ------------------------------------------------------------------------------------------
    enum Ticket : int
    {
        TicketCircles = 0,
        TicketIntersections,
        TicketCellScan,
        TicketCircleScan,
        TicketTorus,
        TicketSdf,
        TicketComp,
        TicketCount = 8
    };
    enum Overflow : uint32_t
    {
        OverflowSectors        = 1u << 0,
        OverflowTrimmed        = 1u << 1,
        OverflowFullCircles    = 1u << 2,
        OverflowIntersections  = 1u << 3,
        OverflowAtomRecords    = 1u << 4,
        OverflowCircleLists    = 1u << 5,
        OverflowSegments       = 1u << 6,
        OverflowTorus          = 1u << 7,
        OverflowProbeNeighbors = 1u << 8,
        OverflowGrid           = 1u << 9, // an atom lies outside its molecule's cell brick (moved atoms)
    };
    struct Control
    {
        uint32_t epoch;                 // incremented per build: tags look-back flags
        uint32_t tickets[ TicketCount ];
        uint32_t radiusMaxBits;         // max vdW radius (float bits, non-negative)
        uint32_t overflow;              // Overflow bits (capacity exceeded: results incomplete)
        uint32_t nbOverflowAtoms;       // atoms that hit the neighbour capacity (truncated, as the reference)
        uint32_t sectorNb;              // V
        uint32_t intersectedCircleNb;   // Ci (complete surface: status 1, i < j)
        uint32_t fullCircleNb;          // Cf
        uint32_t visibleCircleNb;       // exterior: status != 0 among i < j
        uint32_t intersectionWithNbNb;  // Iw
        uint32_t intersectionNoNbNb;    // I - Iw
        uint32_t circleListNb;          // sum of per-circle intersection counts (= 3 I)
        uint32_t segmentNb;             // G
        uint32_t torusNb;               // T
        uint32_t jkAnomalyNb;           // concave records whose jk circle is not intersected
        uint32_t probeNbOverflowNb;     // probes with >= maxIntersectionNeighbors neighbours
        uint32_t circleIntOverflowNb;   // circles with > 15 intersections (segments truncated)
        uint32_t fp64ResolveNb;         // triplets re-solved in FP64
        uint32_t maxAtomRecords;        // largest per-atom concave record count
        uint32_t neighborNb;            // sum of neighbour counts (ordered circles)
        uint32_t nbMaxCount;            // largest neighbour count of a truncated list (before truncation)
        uint32_t pad[ 3 ];
    };

    // ------------------------------------------------------------------------------------------
    // Ordered decoupled look-back: items (atoms, tiles) are processed in ticket order; item t
    // publishes its local counts, walks back over predecessors to obtain its exclusive prefix,
    // then publishes the inclusive value. Flag word: [63:62] state, [61:32] tag, [31:0] value.
    // The tag (build epoch x stage) makes stale flags of previous builds read as "not ready", so
    // flag arrays never need clearing.
    // ------------------------------------------------------------------------------------------
    constexpr uint64_t FlagAggregate = 1ull << 62;
    constexpr uint64_t FlagInclusive = 2ull << 62;

    __device__ __forceinline__ uint64_t pack_flag( uint64_t state, uint32_t tag, uint32_t value )
    {
        return state | ( uint64_t( tag & 0x3fffffffu ) << 32 ) | value;
    }
    __device__ __forceinline__ void flag_store( uint64_t * p, uint64_t v )
    {
        __hip_atomic_store( p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT );
    }
    __device__ __forceinline__ uint64_t flag_load( const uint64_t * p )
    {
        return __hip_atomic_load( const_cast<uint64_t *>( p ), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT );
    }

    // Called by a whole wave with identical t / local[] in every lane. flags holds NC arrays of
    // `stride` words each. Returns the exclusive prefixes in excl[] (identical in every lane).
    template<int NC>
    __device__ void wave_lookback( uint64_t * flags, uint32_t stride, uint32_t t, uint32_t tag, const uint32_t ( &local )[ NC ], uint32_t ( &excl )[ NC ] )
    {
        const int lane = lane_id();
        if ( lane == 0 )
            for ( int c = 0; c < NC; c++ )
                flag_store( flags + size_t( c ) * stride + t, pack_flag( t == 0 ? FlagInclusive : FlagAggregate, tag, local[ c ] ) );
        for ( int c = 0; c < NC; c++ )
            excl[ c ] = 0;
        if ( t == 0 )
            return;
        const uint32_t tag30 = tag & 0x3fffffffu;
        for ( int c = 0; c < NC; c++ )
        {
            const uint64_t * f   = flags + size_t( c ) * stride;
            int64_t          pos = int64_t( t ) - 1;
            uint32_t         acc = 0;
            for ( ;; )
            {
                const int64_t q   = pos - lane;
                uint64_t      w   = q >= 0 ? flag_load( f + q ) : pack_flag( FlagInclusive, tag30, 0 );
                uint32_t      st  = uint32_t( w >> 62 );
                if ( uint32_t( ( w >> 32 ) & 0x3fffffffu ) != tag30 )
                    st = 0;
                const uint64_t incMask = ballot( st == 2 );
                const uint64_t badMask = ballot( st == 0 );
                const int      upto    = incMask ? __ffsll( (unsigned long long)incMask ) - 1 : WaveSize - 1;
                const uint64_t window  = upto == 63 ? ~0ull : ( ( 1ull << ( upto + 1 ) ) - 1 );
                if ( badMask & window )
                {
                    __builtin_amdgcn_s_sleep( 1 );
                    continue; // a predecessor has not published yet
                }
                const uint32_t v = lane <= upto ? uint32_t( w ) : 0u;
                acc += wave_sum( v );
                if ( incMask )
                    break;
                pos -= WaveSize;
            }
            excl[ c ] = acc;
        }
        if ( lane == 0 )
            for ( int c = 0; c < NC; c++ )
                flag_store( flags + size_t( c ) * stride + t, pack_flag( FlagInclusive, tag, excl[ c ] + local[ c ] ) );
    }

    // As wave_lookback, with the columns walking back together: each round reads one window of 64
    // predecessors of every column still open, so the columns share the latency of each round.
    template<int NC>
    __device__ void wave_lookback_joint( uint64_t * flags, uint32_t stride, uint32_t t, uint32_t tag, const uint32_t ( &local )[ NC ], uint32_t ( &excl )[ NC ] )
    {
        const int lane = lane_id();
        auto      mine = [ & ]( const uint32_t( &v )[ NC ] ) { // column `lane` of v
            uint32_t x = v[ 0 ];
#pragma unroll
            for ( int c = 1; c < NC; c++ )
                x = lane == c ? v[ c ] : x;
            return x;
        };
        if ( lane < NC )
            flag_store( flags + size_t( lane ) * stride + t, pack_flag( t == 0 ? FlagInclusive : FlagAggregate, tag, mine( local ) ) );
#pragma unroll
        for ( int c = 0; c < NC; c++ )
            excl[ c ] = 0;
        if ( t == 0 )
            return;
        const uint32_t tag30 = tag & 0x3fffffffu;
        bool           done[ NC ];
#pragma unroll
        for ( int c = 0; c < NC; c++ )
            done[ c ] = false;
        for ( int64_t pos = int64_t( t ) - 1;; )
        {
            const int64_t q = pos - lane;
            uint64_t      wv[ NC ];
            int           upto[ NC ];
            bool          fin[ NC ], wait = false;
#pragma unroll
            for ( int c = 0; c < NC; c++ )
            {
                wv[ c ] = 0, upto[ c ] = -1, fin[ c ] = true;
                if ( done[ c ] ) // wave-uniform
                    continue;
                wv[ c ] = q >= 0 ? flag_load( flags + size_t( c ) * stride + q ) : pack_flag( FlagInclusive, tag30, 0 );
            }
#pragma unroll
            for ( int c = 0; c < NC; c++ )
            {
                if ( done[ c ] )
                    continue;
                uint32_t st = uint32_t( wv[ c ] >> 62 );
                if ( uint32_t( ( wv[ c ] >> 32 ) & 0x3fffffffu ) != tag30 )
                    st = 0;
                const uint64_t incMask = ballot( st == 2 );
                const uint64_t badMask = ballot( st == 0 );
                upto[ c ]              = incMask ? __ffsll( (unsigned long long)incMask ) - 1 : WaveSize - 1;
                const uint64_t window  = upto[ c ] == 63 ? ~0ull : ( ( 1ull << ( upto[ c ] + 1 ) ) - 1 );
                wait                   = wait || ( badMask & window ) != 0;
                fin[ c ]               = incMask != 0;
            }
            if ( wait )
            {
                __builtin_amdgcn_s_sleep( 1 );
                continue; // a predecessor has not published yet: read the window again
            }
            bool all = true;
#pragma unroll
            for ( int c = 0; c < NC; c++ )
            {
                if ( done[ c ] )
                    continue;
                excl[ c ] += wave_sum( lane <= upto[ c ] ? uint32_t( wv[ c ] ) : 0u );
                done[ c ] = fin[ c ];
                all       = all && fin[ c ];
            }
            if ( all )
                break;
            pos -= WaveSize;
        }
        uint32_t incl[ NC ];
#pragma unroll
        for ( int c = 0; c < NC; c++ )
            incl[ c ] = excl[ c ] + local[ c ];
        if ( lane < NC )
            flag_store( flags + size_t( lane ) * stride + t, pack_flag( FlagInclusive, tag, mine( incl ) ) );
    }

    // Wave-uniform ticket: lane 0 draws, the wave shares it.
    __device__ __forceinline__ uint32_t wave_ticket( uint32_t * counter )
    {
        uint32_t t = 0;
        if ( lane_id() == 0 )
            t = atomicAdd( counter, 1u );
        return wave_bcast( t, 0 );
    }

    __device__ __forceinline__ void set_overflow( Control * ctl, uint32_t bit ) { atomicOr( &ctl->overflow, bit ); }

    // Look-back tag of one stage of the current build.
    __device__ __forceinline__ uint32_t stage_tag( const Control * ctl, uint32_t stage ) { return ctl->epoch * 8u + stage; }

    // Orders this wave's LDS writes before its later LDS reads of other lanes' values.
    // Workgroups are dealt to the 8 XCDs round-robin (block b runs on XCD b % 8), each XCD with
    // its own L2. Remapping so that each XCD takes a contiguous range of blocks keeps spatially
    // close atoms (sorted by cell) and their neighbours in one L2. Blocks past the last full round
    // keep their index; the mapping is a bijection either way.
    constexpr uint32_t XcdCount = 8;
    __device__ __forceinline__ uint32_t xcd_block()
    {
        const uint32_t b = blockIdx.x, g = gridDim.x, full = ( g / XcdCount ) * XcdCount;
        return b < full ? ( b % XcdCount ) * ( g / XcdCount ) + b / XcdCount : b;
    }

    __device__ __forceinline__ void wave_sync()
    {
        __builtin_amdgcn_fence( __ATOMIC_SEQ_CST, "wavefront" );
        __builtin_amdgcn_wave_barrier();
    }

    // ------------------------------------------------------------------------------------------
    // Molecule table (one entry per molecule of a batch; a single build is a batch of one).
    // Atoms of molecule m occupy [atomBase, atomBase + atomNb) both in input and sorted order;
    // its cells occupy [cellBase, cellBase + dx*dy*dz) of the global cell table.
    // ------------------------------------------------------------------------------------------
    struct Molecule
    {
        float    ox, oy, oz;    // grid origin
        float    cell;          // cell edge (A)
        int      dx, dy, dz;    // cells per axis
        uint32_t cellBase;
        uint32_t atomBase;
        uint32_t atomNb;
        uint32_t pad[ 2 ];
    };

    __device__ __forceinline__ uint32_t find_molecule( const Molecule * mols, uint32_t molNb, uint32_t atom )
    {
        uint32_t lo = 0, hi = molNb; // last m with atomBase <= atom
        while ( hi - lo > 1 )
        {
            const uint32_t mid = ( lo + hi ) >> 1;
            if ( mols[ mid ].atomBase <= atom )
                lo = mid;
            else
                hi = mid;
        }
        return lo;
    }

    __device__ __forceinline__ int clampi( int v, int lo, int hi ) { return v < lo ? lo : ( v > hi ? hi : v ); }

    __device__ __forceinline__ void cell_of( const Molecule & m, float x, float y, float z, int & gx, int & gy, int & gz )
    {
        gx = clampi( int( floorf( ( x - m.ox ) / m.cell ) ), 0, m.dx - 1 );
        gy = clampi( int( floorf( ( y - m.oy ) / m.cell ) ), 0, m.dy - 1 );
        gz = clampi( int( floorf( ( z - m.oz ) / m.cell ) ), 0, m.dz - 1 );
    }
    // Gap between coordinate v and the cell slab [lo, lo + cell] along one axis (0 inside).
    __device__ __forceinline__ float slab_gap( float lo, float cell, float v ) { return fmaxf( fmaxf( lo - v, v - ( lo + cell ) ), 0.f ); }

    __device__ __forceinline__ uint32_t cell_index( const Molecule & m, int gx, int gy, int gz )
    {
        return m.cellBase + uint32_t( ( gz * m.dy + gy ) * m.dx + gx );
    }
} // namespace sesdf
