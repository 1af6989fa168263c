// Contour-Buildup comparator on gfx950: the reference's second surface method, built here only so
// that the SES builder can be timed against it on the same GPU and inputs.
//
// It replaces bcs::ContourBuildup (sesdf/src/contour_buildup/contour_buildup.cu:400-660) and the
// MegaMol kernels it launches (sesdf/extern/ContourBuildup/src/particles_kernel.cu:1525-2381):
// neighbours on a 16^3 grid over the molecule's box, small circles, removal of covered circles,
// the arcs left on each circle (the ten-case interval update), probe positions, spherical
// triangles, torus records and the singularity texture. Results follow the reference's algorithm
// on the same inputs, including its stencil of cells [g - c, g + c) per axis and its capacities
// (64 neighbours per atom, 32 neighbours per probe).
//
// MI355X design (none of it follows the reference's launch shapes):
//   * one wave per atom, one lane per neighbour slot: the reference's capacity of 64 neighbours is
//     the wave width, so atom i's neighbour list, small circles, covered-circle flags and arc lists
//     live in registers and LDS for the whole pipeline; no N x 64 x 64 arc array (6.5 GB at 100k
//     atoms in the reference) and no N x 64 circle arrays are written to HBM;
//   * the neighbour search gathers 64 stencil cells per step and walks only the non-empty ones,
//     testing 64 candidates per step with a ballot-ordered append (the reference's scan order);
//   * covered circles become a 64-bit wave mask, so the arc pass skips them uniformly;
//   * probe and torus offsets come from an ordered decoupled look-back over atoms (tickets), so the
//     arcs are written once, straight to their final place, in the reference's order;
//   * probes are sorted by cell with the library's own stable radix sort.
// Each build reads the counts back once (the reference blocks on five copies).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "common.h"
#include "handle.h"
#include "sesdf_c.h"
#include "sort.h"

namespace sesdf
{
    namespace cb
    {
        constexpr uint32_t K         = 64; // atomNeighborCount (contour_buildup.cu:57)
        constexpr uint32_t ProbeK    = 32; // probeNeighborCount (contour_buildup.cu:57)
        constexpr uint32_t Sentinel  = 0xffu; // arc end of the initial full circle: atom index 0
        constexpr float    TwoPiCb   = 6.28318530718f; // the constant of particles_kernel.cu:1818
        constexpr uint32_t ArcCapSmall = 8;  // arcs per lane in LDS (the reference's largest lists
                                             // on proteins and globules hold 6)
        constexpr uint32_t ArcCapFull  = 80; // the reference's 64 arcs + 16 new ones

        enum CbOverflow : uint32_t
        {
            OverflowProbes = 1u << 0,
            OverflowTori   = 1u << 1,
            OverflowArcs   = 1u << 2,
            OverflowStage  = 1u << 3,
        };

        struct Ctl
        {
            uint32_t epoch;
            uint32_t ticket;
            uint32_t probeNb;     // probes written (min(need, capacity))
            uint32_t torusNb;
            uint32_t probeNeed;   // probes of the build
            uint32_t torusNeed;
            uint32_t overflow;    // CbOverflow bits
            uint32_t nbOverflow;  // atoms with more than 64 neighbours (lists truncated, as the reference)
            uint32_t noContour;   // atoms whose contour vanished
            uint32_t stageNeed;   // largest per-atom probe count beyond the staging stride
            uint32_t pad[ 6 ];
        };

        struct Params
        {
            float    ox, oy, oz;  // worldOrigin (aabb.min)
            float    cx, cy, cz;  // cellSize (aabb extent / dim)
            int      dim;         // cells per axis (power of two)
            float    rp;
            uint32_t texSize;
        };

        __device__ __forceinline__ void grid_pos( const Params & p, float x, float y, float z, int & gx, int & gy, int & gz )
        {
            gx = int( floorf( ( x - p.ox ) / p.cx ) );
            gy = int( floorf( ( y - p.oy ) / p.cy ) );
            gz = int( floorf( ( z - p.oz ) / p.cz ) );
        }
        __device__ __forceinline__ uint32_t grid_hash( const Params & p, int gx, int gy, int gz )
        {
            const int m = p.dim - 1;
            return ( uint32_t( gz & m ) * uint32_t( p.dim ) + uint32_t( gy & m ) ) * uint32_t( p.dim ) + uint32_t( gx & m );
        }

        // sqrtf(d2) < r, deciding without the square root when d2 is farther than 1e-6 (relative)
        // from r^2: the correctly rounded root then lies on the same side of r. NaN compares false.
        __device__ __forceinline__ bool closer( float d2, float r )
        {
            const float r2 = r * r;
            if ( d2 < r2 * ( 1.0f - 1e-6f ) )
                return true;
            if ( !( d2 <= r2 * ( 1.0f + 1e-6f ) ) )
                return false;
            return sqrtf( d2 ) < r;
        }

        __global__ void begin_kernel( Ctl * c )
        {
            uint32_t * w = reinterpret_cast<uint32_t *>( c );
            if ( threadIdx.x == 0 )
                w[ 0 ] = w[ 0 ] + 1; // epoch: tags this build's look-back flags
            else if ( threadIdx.x < sizeof( Ctl ) / 4 )
                w[ threadIdx.x ] = 0;
        }

        // Cell hash and index of every point (calcHashD, particles_kernel.cu:86-107). Slots past the
        // live count (probes: read from the device) get the largest key, so the stable sort leaves
        // them after every live point.
        __global__ void hash_kernel( const float4 * pts, uint32_t cap, const uint32_t * countPtr, Params p, uint32_t * keys,
                                     uint32_t * vals )
        {
            const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
            if ( i >= cap )
                return;
            const uint32_t n = countPtr ? *countPtr : cap;
            uint32_t       h = 0xffffffffu;
            if ( i < n )
            {
                const float4 a = pts[ i ];
                int          gx, gy, gz;
                grid_pos( p, a.x, a.y, a.z, gx, gy, gz );
                h = grid_hash( p, gx, gy, gz );
            }
            keys[ i ] = h;
            vals[ i ] = i;
        }

        // Gather into sorted order (reorderDataAndFindCellStartD, :110-176).
        __global__ void gather_kernel( const uint32_t * vals, const float4 * src, uint32_t cap, const uint32_t * countPtr,
                                       float4 * sorted, uint32_t * sortedToInitial )
        {
            const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
            const uint32_t n = countPtr ? *countPtr : cap;
            if ( s >= n )
                return;
            const uint32_t v     = vals[ s ];
            sorted[ s ]          = src[ v ];
            sortedToInitial[ s ] = v;
        }

        // First sorted point of every cell, cells + 1 entries (a cell's points are
        // [first[h], first[h + 1])): the reference's cellStart / cellEnd, without empty markers, so a
        // run of consecutive cells is read with two loads.
        __global__ void cell_first_kernel( const uint32_t * keys, uint32_t cap, const uint32_t * countPtr, uint32_t cells,
                                           uint32_t * first )
        {
            const uint32_t h = blockIdx.x * blockDim.x + threadIdx.x;
            if ( h > cells )
                return;
            uint32_t lo = 0, hi = countPtr ? *countPtr : cap; // first key >= h
            while ( lo < hi )
            {
                const uint32_t mid = ( lo + hi ) >> 1;
                if ( keys[ mid ] < h )
                    lo = mid + 1;
                else
                    hi = mid;
            }
            first[ h ] = lo;
        }

        struct AtomArgs
        {
            Ctl *            ctl;
            const float4 *   atoms;      // sorted
            const uint32_t * cellFirst;  // cells + 1: first sorted atom of each cell (of any later cell if empty)
            uint32_t *       nbCount;    // N: neighbours after the covered-circle pass
            uint32_t *       nbIds;      // N * 64
            uint32_t *       atomProbes; // N: probes written from atom i
            uint32_t *       atomTori;   // N: tori of atom i
            uint64_t *       visible;    // N: neighbour slots whose circle is a torus
            float4 *         stage;      // N * stageStride: probe xyz, k | slot << 24
            uint32_t         stageStride;
            uint32_t         n;
            Params           prm;
        };

        template<int CAP>
        struct AtomLds
        {
            float4   v[ K ];                           // small circle of slot s: centre vector, |v|^2
            float4   nrm[ K ];                         // unit vector from the neighbour to atom i; w: the
                                                       // neighbour's sorted index (bits)
            float    s[ CAP ][ K ], e[ CAP ][ K ];     // arcs of the lane's circle: angles
            uint32_t key[ CAP ][ K ];                  // start slot | end slot << 8 | valid << 16
        };

        __device__ __forceinline__ float fmod2pi( float x ) { return fmodf( x, TwoPiCb ); }

        // Point of the line where the planes of two small circles meet that is closest to atom i's
        // centre (each plane holds its circle's centre vector v and is normal to it):
        // h = (vj (vj.(vj - vk)) |vk|^2 + vk ((vk - vj).vk) |vj|^2) / (|vj|^2 |vk|^2 - (vj.vk)^2)
        // (particles_kernel.cu:1731-1732), with one hardware reciprocal of the denominator. The
        // reference compiles with fast math (approximate division); either way the decisions that
        // follow differ from exact arithmetic only inside the oracle's epsilon band.
        __device__ __forceinline__ f3 plane_point( f3 vj, f3 vk, float vjvj, float vkvk )
        {
            const float vjvk = dot( vj, vk );
            const float inv  = frcp( vjvj * vkvk - vjvk * vjvk );
            const float c1   = dot( vj, vj - vk ) * vkvk * inv;
            const float c2   = dot( vk - vj, vk ) * vjvj * inv;
            return vj * c1 + vk * c2;
        }

        // One wave per atom (cell-sorted order), lane s = neighbour slot s.
        template<int CAP, int WAVES>
        __global__ void __launch_bounds__( WaveSize * WAVES ) atoms_kernel( AtomArgs a )
        {
            __shared__ AtomLds<CAP> ldsAll[ WAVES ];
            AtomLds<CAP> &           L    = ldsAll[ wave_in_block() ];
            const int                lane = lane_id();
            // consecutive (cell-sorted) atoms per XCD: their neighbours share that XCD's L2
            const uint32_t           i    = xcd_block() * uint32_t( WAVES ) + uint32_t( wave_in_block() );
            if ( i >= a.n )
                return;
            const float4   ai = a.atoms[ i ];
            const Params & p  = a.prm;
            const f3       pi = xyz( ai );

            // ---- neighbours and small circles (findNeighborsCBCuda) ----
            // The stencil's cells of one (y, z) row are consecutive in the hash, so the row's atoms
            // are one contiguous run of sorted indices; the runs of up to 64 rows are numbered as one
            // stream (row order, then index: the reference's scan order) and tested 64 at a time.
            int gx, gy, gz;
            grid_pos( p, ai.x, ai.y, ai.z, gx, gy, gz );
            const float range = ai.w + 3.0f + 2.0f * p.rp;
            const int   cx = int( ceilf( range / p.cx ) ), cy = int( ceilf( range / p.cy ) ), cz = int( ceilf( range / p.cz ) );
            const int   x0 = max( gx - cx, 0 ), x1 = min( gx + cx, p.dim );
            const int   y0 = max( gy - cy, 0 ), y1 = min( gy + cy, p.dim );
            const int   z0 = max( gz - cz, 0 ), z1 = min( gz + cz, p.dim );
            const int   ey = max( y1 - y0, 0 ), ez = max( z1 - z0, 0 );
            const uint32_t rows  = x1 > x0 ? uint32_t( ey ) * uint32_t( ez ) : 0u;
            const float    Ri    = ai.w + p.rp;
            uint32_t       count = 0;
            bool           full  = false;
            for ( uint32_t r0 = 0; r0 < rows && !full; r0 += WaveSize )
            {
                const uint32_t r  = r0 + uint32_t( lane );
                uint32_t       rs = 0, rn = 0;
                if ( r < rows )
                {
                    const int y = y0 + int( r % uint32_t( ey ) ), z = z0 + int( r / uint32_t( ey ) );
                    rs          = a.cellFirst[ grid_hash( p, x0, y, z ) ];
                    rn          = a.cellFirst[ grid_hash( p, x1 - 1, y, z ) + 1 ] - rs;
                }
                const uint32_t incl  = wave_incl_scan( rn );
                const uint32_t total = wave_bcast( incl, WaveSize - 1 );
                // stream position u -> sorted atom: the run holding u is the number of runs ending at
                // or before it (binary search over the lanes' inclusive ends)
                const uint32_t nr    = min( rows - r0, uint32_t( WaveSize ) );
                auto           atom_at = [ & ]( uint32_t u ) {
                    int q = 0;
                    for ( int step = WaveSize / 2; step > 0; step >>= 1 )
                        if ( uint32_t( step ) < nr * 2 && __shfl( incl, q + step - 1, WaveSize ) <= u )
                            q += step;
                    return __shfl( rs, q, WaveSize ) + ( u - ( __shfl( incl, q, WaveSize ) - __shfl( rn, q, WaveSize ) ) );
                };
                // loads of the next 64 candidates are issued before the current ones are tested
                uint32_t jn = atom_at( uint32_t( lane ) );
                float4   an = uint32_t( lane ) < total ? a.atoms[ jn ] : make_float4( 0.f, 0.f, 0.f, 0.f );
                for ( uint32_t u0 = 0; u0 < total && !full; u0 += WaveSize )
                {
                    const uint32_t j  = jn;
                    const float4   aj = an;
                    const uint32_t u  = u0 + uint32_t( lane );
                    if ( u0 + WaveSize < total )
                    {
                        jn = atom_at( u + WaveSize );
                        if ( u + WaveSize < total )
                            an = a.atoms[ jn ];
                    }
                    bool pass = false;
                    if ( u < total && j != i )
                    {
                        const f3 rel = xyz( aj ) - pi;
                        pass         = closer( dot( rel, rel ), ai.w + aj.w + 2.0f * p.rp );
                    }
                    const uint64_t m    = ballot( pass );
                    const uint32_t slot = count + rank_below( m );
                    if ( pass && slot < K )
                        L.nrm[ slot ].w = __uint_as_float( j );
                    const uint32_t add = uint32_t( __popcll( m ) );
                    if ( count + add > K )
                    {
                        full  = true;
                        count = K;
                    }
                    else
                        count += add;
                }
            }
            if ( full && lane == 0 )
                atomicAdd( &a.ctl->nbOverflow, 1u );
            wave_sync();

            const uint32_t n0  = count;
            const bool     act = uint32_t( lane ) < n0;
            f3             vj  = mk( 0.f, 0.f, 0.f ), pj = mk( 0.f, 0.f, 0.f );
            float          wj  = 0.f;
            uint32_t       j   = 0;
            if ( act )
            {
                // the small circle of slot `lane`, from the same FP32 expressions the search tested
                j               = __float_as_uint( L.nrm[ lane ].w );
                const float4 aj = a.atoms[ j ];
                const f3     rel  = xyz( aj ) - pi;
                const float  dist = sqrtf( dot( rel, rel ) );
                const float  Rj   = aj.w + p.rp;
                float        r    = Ri * Ri + dist * dist - Rj * Rj;
                r                 = r / ( 2.0f * dist * dist );
                vj                = rel * r;
                wj                = sqrtf( Ri * Ri - dot( vj, vj ) );
                pj                = xyz( aj );
                a.nbIds[ size_t( i ) * K + lane ] = j;
            }

            // ---- covered circles (removeCoveredSmallCirclesCBCuda) ----
            // Per-neighbour terms every lane needs (the same expressions as in the pair loop, so the
            // values are bit-identical) go to LDS once. `meet` collects, per lane, the neighbours k
            // whose plane meets circle j's inside atom i's extended sphere: the arc pass visits
            // exactly those (it skips |h| > R with the same h).
            const f3    nj   = act ? unit( pi - pj ) : mk( 0.f, 0.f, 0.f );
            const float vjvj = dot( vj, vj );
            if ( act )
            {
                L.v[ lane ]   = make_float4( vj.x, vj.y, vj.z, vjvj );
                L.nrm[ lane ] = make_float4( nj.x, nj.y, nj.z, __uint_as_float( j ) );
            }
            // slots whose circle centre lies on the far side of atom i from their atom (the arc
            // pass swaps the two cut points for them)
            const uint64_t flipped = ballot( act && dot( vj, pj - pi ) < 0.0f );
            wave_sync();
            bool     addJ = true, noC = false;
            uint64_t meet = 0;
            for ( uint32_t kk = 0; kk < n0; kk++ )
            {
                if ( !act || kk == uint32_t( lane ) )
                    continue;
                const float4 vv4  = L.v[ kk ];
                const f3    vk    = mk( vv4.x, vv4.y, vv4.z );
                const float vkvk  = vv4.w;
                const f3    hv    = plane_point( vj, vk, vjvj, vkvk );
                const bool  far   = dot( hv, hv ) > Ri * Ri;
                meet |= uint64_t( !far ) << kk;
                const float4 nk4 = L.nrm[ kk ];
                const f3    nk   = mk( nk4.x, nk4.y, nk4.z );
                const f3    q    = vk - vj;
                const float njnk = dot( nj, nk );
                if ( njnk == 1.0f )
                {
                    if ( dot( nj, q ) > 0.0f )
                        addJ = false;
                }
                else if ( far )
                {
                    const f3 mj = vj - hv, mk2 = vk - hv;
                    if ( njnk > 0.0f )
                    {
                        if ( dot( mj, mk2 ) > 0.0f && dot( nj, q ) > 0.0f )
                            addJ = false;
                    }
                    else if ( dot( mj, mk2 ) > 0.0f && dot( nj, q ) < 0.0f )
                        noC = true;
                }
            }
            const uint64_t removed   = ballot( act && !addJ );
            const bool     noContour = ballot( act && noC ) != 0;
            const uint32_t n         = noContour ? 0u : n0;
            if ( lane == 0 )
            {
                a.nbCount[ i ] = n;
                if ( noContour )
                    atomicAdd( &a.ctl->noContour, 1u );
            }

            // ---- arcs of the lane's circle (computeArcsCBCuda) ----
            uint32_t arcCnt = 0, written = 0;
            bool     overflow = false;
            f3       xAxis = mk( 0.f, 0.f, 0.f ), yAxis = mk( 0.f, 0.f, 0.f );
            const bool mine = act && n > 0 && !( ( removed >> lane ) & 1ull );
            if ( mine )
            {
                L.s[ 0 ][ lane ] = 0.0f, L.e[ 0 ][ lane ] = TwoPiCb;
                L.key[ 0 ][ lane ] = Sentinel | ( Sentinel << 8 ) | ( 1u << 16 );
                arcCnt             = 1;
                xAxis              = cross( vj, mk( 0.f, 1.f, 0.f ) );
                if ( dot( xAxis, xAxis ) == 0.0f )
                    xAxis = cross( vj, mk( 1.f, 0.f, 0.f ) );
                xAxis = unit( xAxis );
                yAxis = unit( cross( xAxis, vj ) );
            }
            // the neighbours that cut the lane's circle, in slot order (covered circles cut nothing)
            uint64_t cutters = mine ? meet & ~removed : 0ull;
            while ( cutters && !overflow )
            {
                const uint32_t kk = uint32_t( __ffsll( (unsigned long long)cutters ) - 1 );
                cutters &= cutters - 1;
                const float4 vv4  = L.v[ kk ];
                const f3    vk    = mk( vv4.x, vv4.y, vv4.z );
                const float vkvk  = vv4.w;
                const f3    h     = plane_point( vj, vk, vjvj, vkvk );
                const f3    cr   = cross( vk, vj );
                const float root = sqrtf( ( Ri * Ri - dot( h, h ) ) / dot( cr, cr ) );
                f3          q1 = h + cr * root, q2 = h - cr * root;
                if ( ( flipped >> kk ) & 1ull )
                {
                    const f3 tmp = q1;
                    q1           = q2;
                    q2           = tmp;
                }
                float a1 = atan2f( dot( q1 - vj, yAxis ), dot( q1 - vj, xAxis ) );
                float a2 = atan2f( dot( q2 - vj, yAxis ), dot( q2 - vj, xAxis ) );
                if ( a1 > TwoPiCb )
                {
                    a1 = fmod2pi( a1 );
                    a2 = fmod2pi( a2 );
                }
                if ( a2 < a1 )
                    a2 += TwoPiCb;
                if ( a1 < 0.0f )
                {
                    a1 += TwoPiCb;
                    a2 += TwoPiCb;
                }
                const uint32_t kslot  = kk;
                uint32_t       newCnt = arcCnt;
                auto           append = [ & ]( float st, float en, uint32_t sk, uint32_t ek ) {
                    if ( newCnt >= uint32_t( CAP ) )
                    {
                        overflow = true;
                        return;
                    }
                    if ( en < st )
                        en += TwoPiCb;
                    L.s[ newCnt ][ lane ] = st, L.e[ newCnt ][ lane ] = en;
                    L.key[ newCnt ][ lane ] = sk | ( ek << 8 ) | ( 1u << 16 );
                    newCnt++;
                };
                for ( uint32_t q = 0; q < arcCnt; q++ )
                {
                    uint32_t key = L.key[ q ][ lane ];
                    if ( !( key >> 16 ) )
                        continue;
                    const float    s = L.s[ q ][ lane ], e = L.e[ q ][ lane ];
                    const uint32_t sk = key & 0xffu, ek = ( key >> 8 ) & 0xffu;
                    float          ns = s, ne = e;
                    uint32_t       nsk = sk, nek = ek;
                    bool           valid = true;
                    if ( a1 < s )
                    {
                        if ( ( s - a1 ) > ( a2 - a1 ) )
                        {
                            if ( ( ( s - a1 ) + ( e - s ) ) > TwoPiCb )
                            {
                                if ( ( ( s - a1 ) + ( e - s ) ) < ( TwoPiCb + a2 - a1 ) )
                                {
                                    ns = a1, nsk = kslot, ne = fmod2pi( e ); // case (10)
                                }
                                else
                                {
                                    ns = a1, nsk = kslot, ne = fmod2pi( a2 ), nek = kslot;
                                }
                                if ( ne < ns )
                                    ne += TwoPiCb;
                            }
                            else
                                valid = false; // case (1)
                        }
                        else if ( ( ( s - a1 ) + ( e - s ) ) > ( a2 - a1 ) )
                        {
                            ne = fmod2pi( a2 ), nek = kslot; // case (5)
                            if ( ne < ns )
                                ne += TwoPiCb;
                            if ( ( ( s - a1 ) + ( e - s ) ) > TwoPiCb )
                                append( a1, fmod2pi( e ), kslot, ek ); // case (6)
                        } // case (4): unchanged
                    }
                    else if ( ( a1 - s ) > ( e - s ) )
                    {
                        if ( ( ( a1 - s ) + ( a2 - a1 ) ) > TwoPiCb )
                        {
                            if ( ( ( a1 - s ) + ( a2 - a1 ) ) < ( TwoPiCb + e - s ) )
                            {
                                ne = fmod2pi( a2 ), nek = kslot; // case (9)
                                if ( ne < ns )
                                    ne += TwoPiCb;
                            }
                        }
                        else
                            valid = false; // case (2)
                    }
                    else if ( ( ( a1 - s ) + ( a2 - a1 ) ) > ( e - s ) )
                    {
                        ns = a1, nsk = kslot, ne = fmod2pi( e ); // case (7)
                        if ( ne < ns )
                            ne += TwoPiCb;
                        if ( ( ( a1 - s ) + ( a2 - a1 ) ) > TwoPiCb )
                            append( s, fmod2pi( a2 ), sk, kslot ); // case (8)
                    }
                    else
                    {
                        ns = a1, nsk = kslot, ne = fmod2pi( a2 ), nek = kslot; // case (3)
                        if ( ne < ns )
                            ne += TwoPiCb;
                    }
                    L.s[ q ][ lane ] = ns, L.e[ q ][ lane ] = ne;
                    L.key[ q ][ lane ] = nsk | ( nek << 8 ) | ( uint32_t( valid ) << 16 );
                }
                // keep the valid arcs, in order (the new ones last)
                uint32_t o = 0;
                for ( uint32_t q = 0; q < newCnt; q++ )
                {
                    const uint32_t key = L.key[ q ][ lane ];
                    if ( !( key >> 16 ) )
                        continue;
                    if ( o != q )
                    {
                        L.s[ o ][ lane ] = L.s[ q ][ lane ], L.e[ o ][ lane ] = L.e[ q ][ lane ];
                        L.key[ o ][ lane ] = key;
                    }
                    o++;
                }
                arcCnt = o;
            }
            if ( ballot( overflow ) && lane == 0 )
                atomicOr( &a.ctl->overflow, uint32_t( OverflowArcs ) );

            // merge an arc starting at 0 with one ending at 2 pi
            if ( mine && !overflow )
            {
                int idx0 = -1, idx2 = -1;
                for ( uint32_t q = 0; q < arcCnt; q++ )
                {
                    if ( L.s[ q ][ lane ] < 0.00001f )
                        idx0 = int( q );
                    else if ( fabsf( L.e[ q ][ lane ] - TwoPiCb ) < 0.0001f )
                        idx2 = int( q );
                }
                if ( idx0 >= 0 && idx2 >= 0 )
                {
                    const uint32_t k0 = L.key[ idx0 ][ lane ], k2 = L.key[ idx2 ][ lane ];
                    const float    st = L.s[ idx2 ][ lane ];
                    float          en = fmod2pi( L.e[ idx0 ][ lane ] );
                    if ( en < st )
                        en += TwoPiCb;
                    L.s[ idx0 ][ lane ]   = st;
                    L.e[ idx0 ][ lane ]   = en;
                    L.key[ idx0 ][ lane ] = ( k0 & ~0xffu ) | ( k2 & 0xffu );
                    for ( uint32_t q = uint32_t( idx2 ); q + 1 < arcCnt; q++ )
                    {
                        L.s[ q ][ lane ] = L.s[ q + 1 ][ lane ], L.e[ q ][ lane ] = L.e[ q + 1 ][ lane ];
                        L.key[ q ][ lane ] = L.key[ q + 1 ][ lane ];
                    }
                    arcCnt--;
                }
            }
            // atom index of an arc end (slot of N(i), or the initial arc's 0)
            auto endAtom = [ & ]( uint32_t slot ) { return slot == Sentinel ? 0u : __float_as_uint( L.nrm[ slot ].w ); };
            const bool lower = mine && !overflow && i < j;
            if ( lower )
                for ( uint32_t q = 0; q < arcCnt; q++ )
                {
                    const uint32_t key = L.key[ q ][ lane ];
                    written += uint32_t( j < endAtom( key & 0xffu ) ) + uint32_t( j < endAtom( ( key >> 8 ) & 0xffu ) );
                }
            const bool visible = lower && arcCnt > 0;

            // ---- counts per atom; probes staged in the atom's row ----
            const uint64_t visMask = ballot( visible );
            const uint32_t total   = wave_sum( written );
            const uint32_t before  = wave_incl_scan( written ) - written; // every lane takes part
            if ( lane == 0 )
            {
                a.atomProbes[ i ] = total;
                a.atomTori[ i ]   = uint32_t( __popcll( visMask ) );
                a.visible[ i ]    = visMask;
                if ( total > a.stageStride )
                {
                    atomicOr( &a.ctl->overflow, uint32_t( OverflowStage ) );
                    atomicMax( &a.ctl->stageNeed, total );
                }
            }
            if ( written )
            {
                uint32_t r    = before;
                auto     emit = [ & ]( float ang, uint32_t k ) {
                    if ( r < a.stageStride )
                    {
                        const f3 x = ( pi + vj ) + ( xAxis * cosf( ang ) + yAxis * sinf( ang ) ) * wj;
                        a.stage[ size_t( i ) * a.stageStride + r ] =
                            make_float4( x.x, x.y, x.z, __uint_as_float( k | ( uint32_t( lane ) << 24 ) ) );
                    }
                    r++;
                };
                for ( uint32_t q = 0; q < arcCnt; q++ )
                {
                    const uint32_t key = L.key[ q ][ lane ];
                    const uint32_t ks = endAtom( key & 0xffu ), ke = endAtom( ( key >> 8 ) & 0xffu );
                    if ( j < ks )
                        emit( L.s[ q ][ lane ], ks );
                    if ( j < ke )
                        emit( L.e[ q ][ lane ], ke );
                }
            }
        }

        // Probe count and capacity checks after the scan of per-atom counts.
        __global__ void finish_counts_kernel( Ctl * c, uint32_t capProbes, uint32_t capTori )
        {
            if ( threadIdx.x != 0 )
                return;
            c->probeNb = min( c->probeNeed, capProbes );
            c->torusNb = min( c->torusNeed, capTori );
            if ( c->probeNeed > capProbes )
                c->overflow |= OverflowProbes;
            if ( c->torusNeed > capTori )
                c->overflow |= OverflowTori;
        }

        struct EmitArgs
        {
            const float4 *   atoms;
            const uint32_t * nbIds;
            const uint32_t * atomProbes;
            const uint32_t * probeOff;
            const uint32_t * torusOff;
            const uint64_t * visible;
            const float4 *   stage;
            uint32_t         stageStride;
            float4 *         probePos;
            float4 *         tria1;
            float4 *         tria2;
            float4 *         tria3;
            uint4 *          probeIds;  // (i, j, k, 0) sorted atom indices
            float4 *         torusPos;
            float4 *         torusAxis;
            float4 *         torusVs;
            uint2 *          torusIds;
            uint32_t         n;
            uint32_t         capProbes;
            uint32_t         capTori;
            float            rp;
        };

        // Probes with their spherical-triangle vectors, and torus records (writeProbePositionsCBCuda,
        // particles_kernel.cu:2179-2257), at the scanned offsets: one wave per atom, a lane per
        // staged probe and per torus slot.
        __global__ void __launch_bounds__( 256 ) emit_kernel( EmitArgs a )
        {
            const uint32_t i    = ( blockIdx.x * blockDim.x + threadIdx.x ) / WaveSize;
            const int      lane = lane_id();
            if ( i >= a.n )
                return;
            const float4   ai  = a.atoms[ i ];
            const f3       pi  = xyz( ai );
            const uint32_t cnt = min( a.atomProbes[ i ], a.stageStride );
            const uint32_t off = a.probeOff[ i ];
            for ( uint32_t r = uint32_t( lane ); r < cnt; r += WaveSize )
            {
                const uint32_t o = off + r;
                if ( o >= a.capProbes )
                    break;
                const float4   sp   = a.stage[ size_t( i ) * a.stageStride + r ];
                const uint32_t bits = __float_as_uint( sp.w );
                const uint32_t k = bits & 0xffffffu, j = a.nbIds[ size_t( i ) * K + ( bits >> 24 ) ];
                const float4   aj = a.atoms[ j ], ak = a.atoms[ k ];
                a.probePos[ o ] = make_float4( sp.x, sp.y, sp.z, a.rp );
                a.tria1[ o ]    = make_float4( ai.x - sp.x, ai.y - sp.y, ai.z - sp.z, 1.0f );
                a.tria2[ o ]    = make_float4( aj.x - sp.x, aj.y - sp.y, aj.z - sp.z, 1.0f );
                a.tria3[ o ]    = make_float4( ak.x - sp.x, ak.y - sp.y, ak.z - sp.z, a.rp * a.rp );
                a.probeIds[ o ] = make_uint4( i, j, k, 0u );
            }
            const uint64_t vis = a.visible[ i ];
            if ( !( ( vis >> lane ) & 1ull ) )
                return;
            const uint32_t o = a.torusOff[ i ] + rank_below( vis );
            if ( o >= a.capTori )
                return;
            // the small circle of slot `lane` again, by the atom kernel's expressions
            const uint32_t j    = a.nbIds[ size_t( i ) * K + lane ];
            const float4   aj   = a.atoms[ j ];
            const f3       pj   = xyz( aj );
            const f3       rel  = pj - pi;
            const float    dist = sqrtf( dot( rel, rel ) );
            const float    Ri = ai.w + a.rp, Rj = aj.w + a.rp;
            float          r  = Ri * Ri + dist * dist - Rj * Rj;
            r                 = r / ( 2.0f * dist * dist );
            const f3    vj    = rel * r;
            const float wj    = sqrtf( Ri * Ri - dot( vj, vj ) );
            const f3    ta    = unit( vj );
            const f3    tc    = vj + pi;
            const f3    ortho = unit( cross( ta, mk( 0.f, 0.f, 1.f ) ) );
            const f3    P     = tc + ortho * wj;
            const f3    X     = unit( P - pi ) * ai.w;
            const float li = sqrtf( dot( P - pi, P - pi ) ), lj = sqrtf( dot( P - pj, P - pj ) );
            f3          C    = ( pj - pi ) * ( li / ( lj + li ) );
            const f3    XC   = X - C;
            const float dd   = sqrtf( dot( XC, XC ) );
            C                = ( C + pi ) - tc;
            a.torusPos[ o ]  = make_float4( tc.x, tc.y, tc.z, wj );
            a.torusAxis[ o ] = make_float4( ta.x, ta.y, ta.z, a.rp );
            a.torusVs[ o ]   = make_float4( C.x, C.y, C.z, dd );
            a.torusIds[ o ]  = make_uint2( i, j );
        }
    } // namespace cb
} // namespace sesdf

namespace sesdf
{
    namespace cb
    {
        // Singularity texture (findProbeNeighborsCBCuda, particles_kernel.cu:2318-2379): per sorted
        // probe, the first 32 other probes closer than 2 rp in scan order, and the texture
        // coordinates of its row. Probes are sorted by cell, so a wave's 64 probes usually share one
        // cell and stencil; the stencil's (y, z) rows are contiguous runs of sorted probes. The wave
        // stages 64 candidates of a run per step in LDS, one per lane, and every lane tests them all
        // in the scan order (LDS broadcast reads, eight at a time); a wave whose probes sit in two
        // cells makes two such passes, and probes beyond those go lane by lane (waves spread over
        // many small cells, as on small molecules). The distance test decides from the squared
        // distance unless it lies within 1e-5 (relative) of (2 rp)^2, where it takes the reference's
        // sqrtf(dx*dx + dy*dy + dz*dz) < 2 rp.
        constexpr int SingWaves = 4;
        __global__ void __launch_bounds__( WaveSize * SingWaves ) singularity_kernel( const float4 * sorted, const uint32_t * probeSorted,
                                                                                    const uint32_t * cellFirst, const uint32_t * countPtr,
                                                                                    Params p, float * texCoord, float * singTex )
        {
            __shared__ float4 chunkAll[ SingWaves ][ WaveSize ];
            float4 *          chunk = chunkAll[ wave_in_block() ];
            const uint32_t    s     = blockIdx.x * blockDim.x + threadIdx.x;
            const uint32_t    n     = *countPtr;
            const bool        valid = s < n;
            if ( ballot( valid ) == 0 )
                return;
            const float4 a = valid ? sorted[ s ] : make_float4( 0.f, 0.f, 0.f, 0.f );
            int          gx, gy, gz;
            grid_pos( p, a.x, a.y, a.z, gx, gy, gz );
            const float range = p.rp + p.rp;
            const float lo2 = range * range * ( 1.0f - 1e-5f ), hi2 = range * range * ( 1.0f + 1e-5f );
            const int   cx = int( ceilf( range / p.cx ) ), cy = int( ceilf( range / p.cy ) ), cz = int( ceilf( range / p.cz ) );
            uint32_t    count = 0;
            bool        full  = false;
            auto        near  = [ & ]( float4 b ) {
                const float dx = b.x - a.x, dy = b.y - a.y, dz = b.z - a.z;
                const float d2 = __builtin_fmaf( dz, dz, __builtin_fmaf( dy, dy, dx * dx ) );
                if ( d2 < lo2 )
                    return true;
                if ( d2 > hi2 )
                    return false;
                return sqrtf( dx * dx + dy * dy + dz * dz ) < range; // the reference's test, exactly
            };
            auto record = [ & ]( float4 b ) {
                if ( count >= ProbeK )
                {
                    full = true;
                    return;
                }
                float * o = singTex + ( size_t( s ) * ProbeK + count ) * 3;
                o[ 0 ] = b.x, o[ 1 ] = b.y, o[ 2 ] = b.z;
                count++;
            };
            uint64_t todo = ballot( valid );
            // at most two cooperative passes (the cells of the first probes still open)
            for ( int pass = 0; todo && pass < 2; pass++ )
            {
                const int  l    = __ffsll( (unsigned long long)todo ) - 1;
                const int  gx0 = wave_bcast( gx, l ), gy0 = wave_bcast( gy, l ), gz0 = wave_bcast( gz, l );
                const bool mine = valid && gx == gx0 && gy == gy0 && gz == gz0;
                todo &= ~ballot( mine );
                const int x0 = max( gx0 - cx, 0 ), x1 = min( gx0 + cx, p.dim );
                if ( x1 <= x0 )
                    continue;
                for ( int z = max( gz0 - cz, 0 ); z < min( gz0 + cz, p.dim ); z++ )
                    for ( int y = max( gy0 - cy, 0 ); y < min( gy0 + cy, p.dim ); y++ )
                    {
                        const uint32_t st = cellFirst[ grid_hash( p, x0, y, z ) ];
                        const uint32_t en = cellFirst[ grid_hash( p, x1 - 1, y, z ) + 1 ];
                        for ( uint32_t base = st; base < en; base += WaveSize )
                        {
                            const uint32_t t = base + uint32_t( lane_id() );
                            chunk[ lane_id() ] = t < en ? sorted[ t ] : make_float4( 0.f, 0.f, 0.f, 0.f );
                            wave_sync();
                            const uint32_t m = min( en - base, uint32_t( WaveSize ) );
                            for ( uint32_t q0 = 0; q0 < m; q0 += 8 )
                            {
                                uint32_t hits = 0;
#pragma unroll
                                for ( uint32_t u = 0; u < 8; u++ )
                                    if ( q0 + u < m && near( chunk[ q0 + u ] ) )
                                        hits |= 1u << u;
                                if ( !mine || full )
                                    continue;
                                while ( hits )
                                {
                                    const uint32_t u = uint32_t( __ffs( hits ) - 1 );
                                    hits &= hits - 1;
                                    if ( base + q0 + u != s && !full )
                                        record( chunk[ q0 + u ] );
                                }
                            }
                            wave_sync();
                        }
                    }
            }
            if ( ( todo >> lane_id() ) & 1ull )
            {
                const int x0 = max( gx - cx, 0 ), x1 = min( gx + cx, p.dim );
                for ( int z = max( gz - cz, 0 ); z < min( gz + cz, p.dim ) && !full && x1 > x0; z++ )
                    for ( int y = max( gy - cy, 0 ); y < min( gy + cy, p.dim ) && !full; y++ )
                    {
                        const uint32_t en = cellFirst[ grid_hash( p, x1 - 1, y, z ) + 1 ];
                        for ( uint32_t t = cellFirst[ grid_hash( p, x0, y, z ) ]; t < en && !full; t++ )
                        {
                            const float4 b = sorted[ t ];
                            if ( t != s && near( b ) )
                                record( b );
                        }
                    }
            }
            if ( !valid )
                return;
            float * tc = texCoord + size_t( probeSorted[ s ] ) * 3;
            tc[ 0 ] = float( count ), tc[ 1 ] = float( s / p.texSize ), tc[ 2 ] = float( s % p.texSize );
        }

        // Search grid of the singularity texture when the reference's cells are large: cells of at
        // least 2 rp over the aabb grown by 10 A (coordinates clamped into it), linear index.
        struct FineGrid
        {
            float ox, oy, oz, cell;
            int   dx, dy, dz;
        };
        __device__ __forceinline__ int fine_axis( float v, float o, float cell, int d )
        {
            return int( fminf( fmaxf( floorf( ( v - o ) / cell ), 0.f ), float( d - 1 ) ) );
        }
        __global__ void fine_hash_kernel( const float4 * pts, uint32_t cap, const uint32_t * countPtr, FineGrid g, uint32_t * keys,
                                          uint32_t * vals )
        {
            const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
            if ( i >= cap )
                return;
            uint32_t h = 0xffffffffu; // past the live count: sorted after every live point
            if ( i < *countPtr )
            {
                const float4 a = pts[ i ];
                h = ( uint32_t( fine_axis( a.z, g.oz, g.cell, g.dz ) ) * uint32_t( g.dy ) + uint32_t( fine_axis( a.y, g.oy, g.cell, g.dy ) ) ) *
                        uint32_t( g.dx ) +
                    uint32_t( fine_axis( a.x, g.ox, g.cell, g.dx ) );
            }
            keys[ i ] = h;
            vals[ i ] = i;
        }

        // Singularity texture through the search grid, row for row what singularity_kernel writes.
        // The reference scans its stencil's cells in hash order, which is the order of the sorted
        // probes, so it records the first 32 (by sorted index) of the probes within 2 rp that lie in
        // its stencil's cells. The 27 search cells around a probe hold every probe within 2 rp (cells
        // of at least 2 rp; clamping moves no two points further apart in cells); of those, the ones
        // whose reference cell (the same wrapped hash) is in the stencil are kept, and an insertion
        // list in LDS (transposed, one column per lane) holds the 32 smallest sorted indices in order.
        // The distance test is singularity_kernel's.
        constexpr int FineWaves = 4;
        __global__ void __launch_bounds__( WaveSize * FineWaves ) singularity_fine_kernel( const float4 * sorted, const uint32_t * probeSorted,
                                                                                         const float4 * finePts, const uint32_t * fineT,
                                                                                         const uint32_t * fineFirst, const uint32_t * countPtr,
                                                                                         Params p, FineGrid g, float * texCoord, float * singTex )
        {
            __shared__ uint32_t lists[ FineWaves ][ ProbeK ][ WaveSize ];
            uint32_t( *lst )[ WaveSize ] = lists[ wave_in_block() ];
            const uint32_t s             = blockIdx.x * blockDim.x + threadIdx.x;
            if ( s >= *countPtr )
                return;
            const int    lane = lane_id();
            const float4 a    = sorted[ s ];
            int          gx, gy, gz;
            grid_pos( p, a.x, a.y, a.z, gx, gy, gz );
            const float range = p.rp + p.rp;
            const float lo2 = range * range * ( 1.0f - 1e-5f ), hi2 = range * range * ( 1.0f + 1e-5f );
            const int   cx = int( ceilf( range / p.cx ) ), cy = int( ceilf( range / p.cy ) ), cz = int( ceilf( range / p.cz ) );
            const int   x0 = max( gx - cx, 0 ), x1 = min( gx + cx, p.dim ), y0 = max( gy - cy, 0 ), y1 = min( gy + cy, p.dim );
            const int   z0 = max( gz - cz, 0 ), z1 = min( gz + cz, p.dim ), m = p.dim - 1;
            uint32_t    count = 0;
            if ( x1 > x0 && y1 > y0 && z1 > z0 )
            {
                const int fx = fine_axis( a.x, g.ox, g.cell, g.dx ), fy = fine_axis( a.y, g.oy, g.cell, g.dy ), fz = fine_axis( a.z, g.oz, g.cell, g.dz );
                const int xa = max( fx - 1, 0 ), xb = min( fx + 1, g.dx - 1 );
                for ( int z = max( fz - 1, 0 ); z <= min( fz + 1, g.dz - 1 ); z++ )
                    for ( int y = max( fy - 1, 0 ); y <= min( fy + 1, g.dy - 1 ); y++ )
                    {
                        const uint32_t row = ( uint32_t( z ) * uint32_t( g.dy ) + uint32_t( y ) ) * uint32_t( g.dx );
                        const uint32_t en  = fineFirst[ row + uint32_t( xb ) + 1 ];
                        for ( uint32_t u = fineFirst[ row + uint32_t( xa ) ]; u < en; u++ )
                        {
                            const uint32_t t = fineT[ u ];
                            if ( t == s || ( count == ProbeK && t > lst[ ProbeK - 1 ][ lane ] ) )
                                continue;
                            const float4 b  = finePts[ u ];
                            const float  dx = b.x - a.x, dy = b.y - a.y, dz = b.z - a.z;
                            const float  d2 = __builtin_fmaf( dz, dz, __builtin_fmaf( dy, dy, dx * dx ) );
                            if ( !( d2 < lo2 ) && ( d2 > hi2 || !( sqrtf( dx * dx + dy * dy + dz * dz ) < range ) ) )
                                continue;
                            int bx, by, bz;
                            grid_pos( p, b.x, b.y, b.z, bx, by, bz );
                            bx &= m, by &= m, bz &= m;
                            if ( bx < x0 || bx >= x1 || by < y0 || by >= y1 || bz < z0 || bz >= z1 )
                                continue;
                            int j = int( min( count, ProbeK - 1 ) );
                            for ( ; j > 0 && lst[ j - 1 ][ lane ] > t; j-- )
                                lst[ j ][ lane ] = lst[ j - 1 ][ lane ];
                            lst[ j ][ lane ] = t;
                            count            = min( count + 1, ProbeK );
                        }
                    }
                for ( uint32_t j = 0; j < count; j++ )
                {
                    const float4 b = sorted[ lst[ j ][ lane ] ];
                    float *      o = singTex + ( size_t( s ) * ProbeK + j ) * 3;
                    o[ 0 ] = b.x, o[ 1 ] = b.y, o[ 2 ] = b.z;
                }
            }
            float * tc = texCoord + size_t( probeSorted[ s ] ) * 3;
            tc[ 0 ] = float( count ), tc[ 1 ] = float( s / p.texSize ), tc[ 2 ] = float( s % p.texSize );
        }

        constexpr int Waves = 4; // waves per workgroup of the atom kernel (small arc lists)

        struct Handle
        {
            int             device    = 0;
            hipStream_t     stream    = nullptr;
            bool            ownStream = false;
            bool            useGraph  = true;
            Params          prm {};
            uint32_t        n      = 0;
            const float *   host   = nullptr;
            uint32_t        cells  = 0;
            int             bits   = 0;
            uint32_t        arcCap = ArcCapSmall;
            uint32_t        capP = 0, capT = 0, capSort = 0;
            Ctl *           ctl  = nullptr;
            Ctl *           hctl = nullptr;
            float4 *        input = nullptr, *atoms = nullptr;
            uint32_t *      sortedToInitial = nullptr, *nbCount = nullptr, *nbIds = nullptr;
            uint32_t *      keys0 = nullptr, *vals0 = nullptr, *keys1 = nullptr, *vals1 = nullptr, *hist = nullptr;
            uint32_t *      cellFirst = nullptr, *pCellFirst = nullptr; // cells + 1 each
            uint32_t *      atomProbes = nullptr, *atomTori = nullptr, *probeOff = nullptr, *torusOff = nullptr;
            uint64_t *      visible = nullptr;
            float4 *        stage   = nullptr;
            uint32_t        stageStride = 32;    // probes staged per atom (largest seen: 22)
            Control *       scanCtl = nullptr;   // tickets and epoch of the count scan
            uint64_t *      scanFlags = nullptr; // its look-back flags
            uint32_t        scanFlagStride = 0;
            float4 *        probePos = nullptr, *tria1 = nullptr, *tria2 = nullptr, *tria3 = nullptr, *sortedProbes = nullptr;
            uint4 *         probeIds = nullptr;
            float4 *        torusPos = nullptr, *torusAxis = nullptr, *torusVs = nullptr;
            uint2 *         torusIds = nullptr;
            uint32_t *      probeSorted = nullptr;
            float *         texCoord = nullptr, *singTex = nullptr;
            // singularity search: 0 by cell size, 1 the reference's stencil scan, 2 the search grid
            int             singSearch = 0;
            bool            fineOn     = false, fineBig = false;
            FineGrid        fine {};
            uint32_t        fineCells = 0;
            int             fineBits  = 0;
            uint32_t *      fineFirst = nullptr, *fineT = nullptr; // fineCells + 1; capP
            float4 *        finePts   = nullptr;                   // capP
            std::vector<void *> fixedAllocs, outAllocs;
            uint64_t        fixedBytes = 0, outBytes = 0;
            hipGraphExec_t  graph     = nullptr;
            uint64_t        serial    = 0;
            uint32_t        attempts  = 0;
            bool            built     = false;

            ~Handle()
            {
                if ( graph )
                    (void)hipGraphExecDestroy( graph );
                for ( void * p : outAllocs )
                    (void)hipFree( p );
                for ( void * p : fixedAllocs )
                    (void)hipFree( p );
                if ( hctl )
                    (void)hipHostFree( hctl );
                if ( ownStream && stream )
                    (void)hipStreamDestroy( stream );
            }
        };

        inline sesdf_status hip_err( hipError_t e, const char * what )
        {
            return report_error( SESDF_ERR_HIP, ( std::string( what ) + ": " + hipGetErrorString( e ) ).c_str() );
        }
#define CB_TRY( expr )                                                                                                           \
    do                                                                                                                           \
    {                                                                                                                            \
        hipError_t _e = ( expr );                                                                                                \
        if ( _e != hipSuccess )                                                                                                  \
            return hip_err( _e, #expr );                                                                                         \
    } while ( 0 )

        template<class T>
        static sesdf_status alloc( std::vector<void *> & list, uint64_t & bytes, T *& p, size_t count, bool zero = false )
        {
            const size_t b = std::max<size_t>( count, 1 ) * sizeof( T );
            void *       q = nullptr;
            hipError_t   e = hipMalloc( &q, b );
            if ( e != hipSuccess )
                return report_error( SESDF_ERR_CAPACITY, "Contour-Buildup: out of device memory" );
            if ( zero )
                CB_TRY( hipMemset( q, 0, b ) );
            list.push_back( q );
            bytes += b;
            p = static_cast<T *>( q );
            return SESDF_OK;
        }
#define CB_ALLOC( list, ... )                                                                                                    \
    do                                                                                                                           \
    {                                                                                                                            \
        sesdf_status _s = alloc( h.list##Allocs, h.list##Bytes, __VA_ARGS__ );                                                   \
        if ( _s != SESDF_OK )                                                                                                    \
            return _s;                                                                                                           \
    } while ( 0 )

        // Buffers sized by the atom count (allocated once per handle).
        static sesdf_status allocate_fixed( Handle & h )
        {
            CB_ALLOC( fixed, h.ctl, 1, true );
            CB_ALLOC( fixed, h.input, h.n );
            CB_ALLOC( fixed, h.atoms, h.n );
            CB_ALLOC( fixed, h.sortedToInitial, h.n );
            CB_ALLOC( fixed, h.nbCount, h.n );
            CB_ALLOC( fixed, h.nbIds, size_t( h.n ) * K );
            CB_ALLOC( fixed, h.cellFirst, size_t( h.cells ) + 1 );
            CB_ALLOC( fixed, h.pCellFirst, size_t( h.cells ) + 1 );
            CB_ALLOC( fixed, h.fineFirst, size_t( h.fineCells ) + 1 );
            CB_ALLOC( fixed, h.atomProbes, h.n );
            CB_ALLOC( fixed, h.atomTori, h.n );
            CB_ALLOC( fixed, h.probeOff, size_t( h.n ) + 1 );
            CB_ALLOC( fixed, h.torusOff, size_t( h.n ) + 1 );
            CB_ALLOC( fixed, h.visible, h.n );
            CB_ALLOC( fixed, h.scanCtl, 1, true );
            h.scanFlagStride = h.n / ScanMinTile + 64;
            CB_ALLOC( fixed, h.scanFlags, size_t( h.scanFlagStride ) * 4, true );
            return SESDF_OK;
        }

        // Output and sort buffers sized by the probe / torus capacities (grown on overflow).
        static sesdf_status allocate_outputs( Handle & h )
        {
            for ( void * p : h.outAllocs )
                CB_TRY( hipFree( p ) );
            h.outAllocs.clear();
            h.outBytes = 0;
            if ( h.graph )
            {
                CB_TRY( hipGraphExecDestroy( h.graph ) );
                h.graph = nullptr;
            }
            h.capSort = std::max( h.n, h.capP );
            CB_ALLOC( out, h.keys0, h.capSort );
            CB_ALLOC( out, h.vals0, h.capSort );
            CB_ALLOC( out, h.keys1, h.capSort );
            CB_ALLOC( out, h.vals1, h.capSort );
            CB_ALLOC( out, h.hist, sort_hist_words( h.capSort ) );
            CB_ALLOC( out, h.stage, size_t( h.n ) * h.stageStride );
            CB_ALLOC( out, h.probePos, h.capP );
            CB_ALLOC( out, h.tria1, h.capP );
            CB_ALLOC( out, h.tria2, h.capP );
            CB_ALLOC( out, h.tria3, h.capP );
            CB_ALLOC( out, h.probeIds, h.capP );
            CB_ALLOC( out, h.sortedProbes, h.capP );
            CB_ALLOC( out, h.probeSorted, h.capP );
            CB_ALLOC( out, h.texCoord, size_t( h.capP ) * 3 );
            CB_ALLOC( out, h.singTex, size_t( h.capP ) * ProbeK * 3 );
            CB_ALLOC( out, h.finePts, h.capP );
            CB_ALLOC( out, h.fineT, h.capP );
            CB_ALLOC( out, h.torusPos, h.capT );
            CB_ALLOC( out, h.torusAxis, h.capT );
            CB_ALLOC( out, h.torusVs, h.capT );
            CB_ALLOC( out, h.torusIds, h.capT );
            return SESDF_OK;
        }

        // Search grid of the singularity texture: cells of 2 rp (0.1% over, and 1e-3 A, so that float
        // rounding never puts two probes within 2 rp two cells apart) over the aabb grown by 10 A, at
        // most 2^24 cells (larger cells past that). It replaces the reference's stencil scan when
        // that stencil holds over 15 times the volume of the 27 search cells: below that, the search
        // grid's own hash, sort and cell ranges cost about what the shorter scan saves (measured on
        // the sample proteins: 1A8R, 7O0U, 3JC8, 7CGO gain; 7DBB and a 100k globule lose).
        static void plan_fine_grid( Handle & h, const float * box )
        {
            const Params & p   = h.prm;
            const float    pad = 10.f;
            float          ext[ 3 ];
            for ( int q = 0; q < 3; q++ )
                ext[ q ] = box[ 3 + q ] - box[ q ] + 2.f * pad;
            float cell = std::max( 2.f * p.rp * 1.001f + 1e-3f, 0.1f );
            int   d[ 3 ];
            for ( ;; )
            {
                uint64_t cells = 1;
                for ( int q = 0; q < 3; q++ )
                    d[ q ] = std::max( 1, int( std::ceil( ext[ q ] / cell ) ) ), cells *= uint64_t( d[ q ] );
                if ( cells <= ( 1ull << 24 ) )
                    break;
                cell *= 1.25f;
            }
            h.fine      = { box[ 0 ] - pad, box[ 1 ] - pad, box[ 2 ] - pad, cell, d[ 0 ], d[ 1 ], d[ 2 ] };
            h.fineCells = uint32_t( d[ 0 ] ) * uint32_t( d[ 1 ] ) * uint32_t( d[ 2 ] );
            h.fineBits  = 0;
            while ( ( 1ull << h.fineBits ) < h.fineCells )
                h.fineBits++;
            const float range = 2.f * p.rp, cs[ 3 ] = { p.cx, p.cy, p.cz };
            double      stencil = 1.0;
            for ( int q = 0; q < 3; q++ )
                stencil *= cs[ q ] > 0.f ? 2.0 * std::ceil( range / cs[ q ] ) * cs[ q ] : 0.0;
            h.fineBig = stencil > 15.0 * 27.0 * double( cell ) * cell * cell;
            h.fineOn  = h.singSearch == 2 || ( h.singSearch == 0 && h.fineBig );
        }

        static uint32_t blocks( uint64_t n, uint32_t per ) { return uint32_t( std::max<uint64_t>( 1, ( n + per - 1 ) / per ) ); }

        // The device stage sequence of one build (captured once into a hipGraph when enabled).
        static void enqueue_stages( Handle & h, hipStream_t s )
        {
            const Params & p = h.prm;
            begin_kernel<<<1, 64, 0, s>>>( h.ctl );
            // atom grid: hash, stable sort by cell, cell ranges and the sorted atoms
            hash_kernel<<<blocks( h.n, 256 ), 256, 0, s>>>( h.input, h.n, nullptr, p, h.keys0, h.vals0 );
            const int w = radix_sort_pairs( h.keys0, h.vals0, h.keys1, h.vals1, h.hist, h.n, h.bits, s );
            cell_first_kernel<<<blocks( uint64_t( h.cells ) + 1, 256 ), 256, 0, s>>>( w ? h.keys1 : h.keys0, h.n, nullptr, h.cells,
                                                                                    h.cellFirst );
            gather_kernel<<<blocks( h.n, 256 ), 256, 0, s>>>( w ? h.vals1 : h.vals0, h.input, h.n, nullptr, h.atoms, h.sortedToInitial );
            // neighbours, covered circles, arcs: counts and staged probes per atom
            AtomArgs a {};
            a.ctl = h.ctl, a.atoms = h.atoms, a.cellFirst = h.cellFirst, a.nbCount = h.nbCount;
            a.nbIds = h.nbIds, a.atomProbes = h.atomProbes, a.atomTori = h.atomTori, a.visible = h.visible;
            a.stage = h.stage, a.stageStride = h.stageStride, a.n = h.n, a.prm = p;
            if ( h.arcCap == ArcCapSmall )
                atoms_kernel<ArcCapSmall, Waves><<<blocks( h.n, Waves ), WaveSize * Waves, 0, s>>>( a );
            else
                atoms_kernel<ArcCapFull, 1><<<h.n, WaveSize, 0, s>>>( a );
            // offsets of every atom's probes and tori (the reference's two exclusive scans)
            Buffers sb {};
            sb.ctl = h.scanCtl, sb.flags = h.scanFlags, sb.flagStride = h.scanFlagStride;
            launch_begin( sb, s );
            ScanArgs sa {};
            sa.in[ 0 ] = h.atomProbes, sa.in[ 1 ] = h.atomTori;
            sa.out[ 0 ] = h.probeOff, sa.out[ 1 ] = h.torusOff;
            sa.totals[ 0 ] = &h.ctl->probeNeed, sa.totals[ 1 ] = &h.ctl->torusNeed;
            sa.n = h.n, sa.ticket = 0, sa.stage = 1;
            launch_scan<2>( sb, sa, s );
            finish_counts_kernel<<<1, 64, 0, s>>>( h.ctl, h.capP, h.capT );
            EmitArgs e {};
            e.atoms = h.atoms, e.nbIds = h.nbIds, e.atomProbes = h.atomProbes, e.probeOff = h.probeOff;
            e.torusOff = h.torusOff, e.visible = h.visible, e.stage = h.stage, e.stageStride = h.stageStride;
            e.probePos = h.probePos, e.tria1 = h.tria1, e.tria2 = h.tria2, e.tria3 = h.tria3, e.probeIds = h.probeIds;
            e.torusPos = h.torusPos, e.torusAxis = h.torusAxis, e.torusVs = h.torusVs, e.torusIds = h.torusIds;
            e.n = h.n, e.capProbes = h.capP, e.capTori = h.capT, e.rp = p.rp;
            emit_kernel<<<blocks( uint64_t( h.n ) * WaveSize, 256 ), 256, 0, s>>>( e );
            // probe grid on the same cells, singularity texture
            hash_kernel<<<blocks( h.capP, 256 ), 256, 0, s>>>( h.probePos, h.capP, &h.ctl->probeNb, p, h.keys0, h.vals0 );
            const int w2 = radix_sort_pairs( h.keys0, h.vals0, h.keys1, h.vals1, h.hist, h.capP, h.bits, s );
            cell_first_kernel<<<blocks( uint64_t( h.cells ) + 1, 256 ), 256, 0, s>>>( w2 ? h.keys1 : h.keys0, h.capP, &h.ctl->probeNb,
                                                                                    h.cells, h.pCellFirst );
            gather_kernel<<<blocks( h.capP, 256 ), 256, 0, s>>>( w2 ? h.vals1 : h.vals0, h.probePos, h.capP, &h.ctl->probeNb,
                                                                 h.sortedProbes, h.probeSorted );
            if ( h.fineOn )
            {
                // sorted probes into the search grid (values: sorted indices), then the texture
                fine_hash_kernel<<<blocks( h.capP, 256 ), 256, 0, s>>>( h.sortedProbes, h.capP, &h.ctl->probeNb, h.fine, h.keys0, h.vals0 );
                const int w3 = radix_sort_pairs( h.keys0, h.vals0, h.keys1, h.vals1, h.hist, h.capP, h.fineBits, s );
                cell_first_kernel<<<blocks( uint64_t( h.fineCells ) + 1, 256 ), 256, 0, s>>>( w3 ? h.keys1 : h.keys0, h.capP, &h.ctl->probeNb,
                                                                                            h.fineCells, h.fineFirst );
                gather_kernel<<<blocks( h.capP, 256 ), 256, 0, s>>>( w3 ? h.vals1 : h.vals0, h.sortedProbes, h.capP, &h.ctl->probeNb,
                                                                     h.finePts, h.fineT );
                singularity_fine_kernel<<<blocks( h.capP, WaveSize * FineWaves ), WaveSize * FineWaves, 0, s>>>(
                    h.sortedProbes, h.probeSorted, h.finePts, h.fineT, h.fineFirst, &h.ctl->probeNb, p, h.fine, h.texCoord, h.singTex );
            }
            else
                singularity_kernel<<<blocks( h.capP, WaveSize * SingWaves ), WaveSize * SingWaves, 0, s>>>(
                    h.sortedProbes, h.probeSorted, h.pCellFirst, &h.ctl->probeNb, p, h.texCoord, h.singTex );
        }

        static sesdf_status enqueue( Handle & h )
        {
            CB_TRY( hipSetDevice( h.device ) );
            h.serial++;
            if ( h.n == 0 )
                return SESDF_OK;
            if ( h.useGraph )
            {
                if ( !h.graph )
                {
                    hipGraph_t g = nullptr;
                    CB_TRY( hipStreamBeginCapture( h.stream, hipStreamCaptureModeThreadLocal ) );
                    enqueue_stages( h, h.stream );
                    CB_TRY( hipStreamEndCapture( h.stream, &g ) );
                    CB_TRY( hipGraphInstantiate( &h.graph, g, nullptr, nullptr, 0 ) );
                    CB_TRY( hipGraphDestroy( g ) );
                }
                CB_TRY( hipGraphLaunch( h.graph, h.stream ) );
            }
            else
                enqueue_stages( h, h.stream );
            CB_TRY( hipGetLastError() );
            return SESDF_OK;
        }

        static sesdf_status readback( Handle & h )
        {
            if ( h.n == 0 )
            {
                std::memset( h.hctl, 0, sizeof( Ctl ) );
                return SESDF_OK;
            }
            CB_TRY( hipMemcpyAsync( h.hctl, h.ctl, sizeof( Ctl ), hipMemcpyDeviceToHost, h.stream ) );
            CB_TRY( hipStreamSynchronize( h.stream ) );
            return SESDF_OK;
        }

        // Builds from the input buffer; grows the capacities (or widens the arc lists) and builds
        // again when the device reports an overflow.
        static sesdf_status build_resident( Handle & h )
        {
            for ( h.attempts = 1;; h.attempts++ )
            {
                sesdf_status st = enqueue( h );
                if ( st != SESDF_OK )
                    return st;
                if ( ( st = readback( h ) ) != SESDF_OK )
                    return st;
                const uint32_t ov = h.hctl->overflow;
                if ( !ov )
                    break;
                if ( h.attempts >= 4 )
                    return report_error( SESDF_ERR_CAPACITY, "Contour-Buildup: capacities still exceeded after 4 attempts" );
                if ( ov & OverflowArcs )
                {
                    if ( h.arcCap == ArcCapFull )
                        return report_error( SESDF_ERR_CAPACITY,
                                             "Contour-Buildup: an arc list exceeds the reference's 64 + 16 slots" );
                    h.arcCap = ArcCapFull;
                    if ( h.graph )
                    {
                        CB_TRY( hipGraphExecDestroy( h.graph ) );
                        h.graph = nullptr;
                    }
                }
                if ( ov & OverflowStage )
                {
                    uint32_t st2 = h.stageStride;
                    while ( st2 < h.hctl->stageNeed )
                        st2 *= 2;
                    h.stageStride = st2;
                }
                if ( ov & ( OverflowProbes | OverflowTori | OverflowStage ) )
                {
                    if ( ov & OverflowProbes )
                        h.capP = uint32_t( std::min<uint64_t>( 0xffffffffull / 8, uint64_t( h.hctl->probeNeed ) * 5 / 4 + 1024 ) );
                    if ( ov & OverflowTori )
                        h.capT = uint32_t( std::min<uint64_t>( 0xffffffffull / 8, uint64_t( h.hctl->torusNeed ) * 5 / 4 + 1024 ) );
                    if ( ( st = allocate_outputs( h ) ) != SESDF_OK )
                        return st;
                }
            }
            h.built = true;
            return SESDF_OK;
        }

        static bool finite_atoms( const float * a, uint32_t n )
        {
            for ( size_t q = 0; q < size_t( n ) * 4; q++ )
                if ( !std::isfinite( a[ q ] ) )
                    return false;
            return true;
        }
    } // namespace cb
} // namespace sesdf

struct sesdf_cb_handle
{
    sesdf::cb::Handle h;
};

using namespace sesdf;
using sesdf::cb::hip_err;

extern "C"
{
    sesdf_status sesdf_cb_create( const sesdf_cb_create_info * info, sesdf_cb_handle ** out )
    {
        if ( !info || !out )
            return report_error( SESDF_ERR_INVALID, "null argument" );
        *out = nullptr;
        const uint32_t dim = info->grid_dimensions ? info->grid_dimensions : 16u;
        if ( dim > 128 || ( dim & ( dim - 1 ) ) )
            return report_error( SESDF_ERR_INVALID, "grid_dimensions must be a power of two, at most 128" );
        if ( info->atom_nb && !info->xyzr )
            return report_error( SESDF_ERR_INVALID, "xyzr is null" );
        if ( info->atom_nb >= ( 1u << 24 ) )
            return report_error( SESDF_ERR_CAPACITY, "at most 2^24 - 1 atoms (arc end atoms are carried as floats)" );
        if ( !( info->probe_radius >= 0.f ) || !std::isfinite( info->probe_radius ) )
            return report_error( SESDF_ERR_INVALID, "probe_radius must be finite and non-negative" );
        if ( info->xyzr && !cb::finite_atoms( info->xyzr, info->atom_nb ) )
            return report_error( SESDF_ERR_INVALID, "atom coordinates and radii must be finite" );
        float box[ 6 ];
        bool  given = false;
        for ( int q = 0; q < 6; q++ )
            given = given || info->aabb[ q ] != 0.f;
        if ( given )
            std::memcpy( box, info->aabb, sizeof( box ) );
        else
        {
            // getAabb (sesdf/src/core/molecule.cpp:132-143): centre -/+ vdW radius
            for ( int q = 0; q < 3; q++ )
                box[ q ] = info->atom_nb ? 3.4e38f : 0.f, box[ 3 + q ] = info->atom_nb ? -3.4e38f : 0.f;
            for ( uint32_t i = 0; i < info->atom_nb; i++ )
                for ( int q = 0; q < 3; q++ )
                {
                    const float c = info->xyzr[ 4 * size_t( i ) + q ], r = info->xyzr[ 4 * size_t( i ) + 3 ];
                    box[ q ]      = std::min( box[ q ], c - r );
                    box[ 3 + q ]  = std::max( box[ 3 + q ], c + r );
                }
        }
        for ( int q = 0; q < 3; q++ )
            if ( !std::isfinite( box[ q ] ) || !std::isfinite( box[ 3 + q ] ) || box[ q ] > box[ 3 + q ] )
                return report_error( SESDF_ERR_INVALID, "aabb must be finite with min <= max" );

        auto *        hh = new sesdf_cb_handle();
        cb::Handle &  h  = hh->h;
        auto fail = [ & ]( sesdf_status st ) {
            delete hh;
            return st;
        };
        h.device   = info->device;
        h.useGraph = info->use_graph != 0;
        h.n        = info->atom_nb;
        h.host     = info->xyzr;
        // worldOrigin = aabb.min, cellSize = extent / gridSize (contour_buildup.cu:76-98)
        cb::Params & p = h.prm;
        p.ox = box[ 0 ], p.oy = box[ 1 ], p.oz = box[ 2 ];
        p.cx = ( box[ 3 ] - box[ 0 ] ) / float( dim ), p.cy = ( box[ 4 ] - box[ 1 ] ) / float( dim ), p.cz = ( box[ 5 ] - box[ 2 ] ) / float( dim );
        p.dim     = int( dim );
        p.rp      = info->probe_radius;
        p.texSize = info->tex_size ? info->tex_size : 32768u;
        h.cells   = dim * dim * dim;
        h.bits    = 0;
        while ( ( 1u << h.bits ) < h.cells )
            h.bits++;
        cb::plan_fine_grid( h, box );
        hipError_t e = hipSetDevice( h.device );
        if ( e != hipSuccess )
            return fail( cb::hip_err( e, "hipSetDevice" ) );
        if ( info->stream )
            h.stream = static_cast<hipStream_t>( info->stream );
        else
        {
            if ( ( e = hipStreamCreateWithFlags( &h.stream, hipStreamNonBlocking ) ) != hipSuccess )
                return fail( cb::hip_err( e, "hipStreamCreate" ) );
            h.ownStream = true;
        }
        if ( ( e = hipHostMalloc( reinterpret_cast<void **>( &h.hctl ), sizeof( cb::Ctl ), hipHostMallocDefault ) ) != hipSuccess )
            return fail( cb::hip_err( e, "hipHostMalloc" ) );
        std::memset( h.hctl, 0, sizeof( cb::Ctl ) );
        h.capP = h.n + h.n / 4 + 1024;
        h.capT = h.n + h.n / 2 + 1024;
        sesdf_status st = cb::allocate_fixed( h );
        if ( st == SESDF_OK )
            st = cb::allocate_outputs( h );
        if ( st != SESDF_OK )
            return fail( st );
        if ( info->build_surface && h.n )
        {
            if ( ( e = hipMemcpyAsync( h.input, h.host, size_t( h.n ) * 16, hipMemcpyHostToDevice, h.stream ) ) != hipSuccess )
                return fail( cb::hip_err( e, "upload" ) );
            if ( ( st = cb::build_resident( h ) ) != SESDF_OK )
                return fail( st );
        }
        *out = hh;
        return SESDF_OK;
    }

    sesdf_status sesdf_cb_destroy( sesdf_cb_handle * h )
    {
        if ( h )
        {
            (void)hipSetDevice( h->h.device );
            (void)hipStreamSynchronize( h->h.stream );
            delete h;
        }
        return SESDF_OK;
    }

    sesdf_status sesdf_cb_set_atoms( sesdf_cb_handle * hh, const float * xyzr )
    {
        if ( !hh || ( !xyzr && hh->h.n ) )
            return report_error( SESDF_ERR_INVALID, "null argument" );
        if ( xyzr && !cb::finite_atoms( xyzr, hh->h.n ) )
            return report_error( SESDF_ERR_INVALID, "atom coordinates and radii must be finite" );
        hh->h.host = xyzr;
        return SESDF_OK;
    }

    sesdf_status sesdf_cb_build( sesdf_cb_handle * hh )
    {
        if ( !hh )
            return report_error( SESDF_ERR_INVALID, "null handle" );
        cb::Handle & h = hh->h;
        CB_TRY( hipSetDevice( h.device ) );
        if ( h.n )
            CB_TRY( hipMemcpyAsync( h.input, h.host, size_t( h.n ) * 16, hipMemcpyHostToDevice, h.stream ) );
        return cb::build_resident( h );
    }

    sesdf_status sesdf_cb_build_device( sesdf_cb_handle * hh, const float * d_xyzr )
    {
        if ( !hh )
            return report_error( SESDF_ERR_INVALID, "null handle" );
        cb::Handle & h = hh->h;
        CB_TRY( hipSetDevice( h.device ) );
        if ( d_xyzr && h.n && d_xyzr != reinterpret_cast<const float *>( h.input ) )
            CB_TRY( hipMemcpyAsync( h.input, d_xyzr, size_t( h.n ) * 16, hipMemcpyDeviceToDevice, h.stream ) );
        return cb::build_resident( h );
    }

    sesdf_status sesdf_cb_input_buffer( sesdf_cb_handle * hh, float ** d_xyzr )
    {
        if ( !hh || !d_xyzr )
            return report_error( SESDF_ERR_INVALID, "null argument" );
        *d_xyzr = reinterpret_cast<float *>( hh->h.input );
        return SESDF_OK;
    }

    sesdf_status sesdf_cb_enqueue( sesdf_cb_handle * hh )
    {
        if ( !hh )
            return report_error( SESDF_ERR_INVALID, "null handle" );
        if ( !hh->h.built && hh->h.n )
            return report_error( SESDF_ERR_NOT_BUILT, "sesdf_cb_enqueue replays a build: run sesdf_cb_build first" );
        return cb::enqueue( hh->h );
    }

    sesdf_status sesdf_cb_sync( sesdf_cb_handle * hh )
    {
        if ( !hh )
            return report_error( SESDF_ERR_INVALID, "null handle" );
        sesdf_status st = cb::readback( hh->h );
        if ( st == SESDF_OK && hh->h.hctl->overflow )
            return report_error( SESDF_ERR_CAPACITY, "Contour-Buildup: a replayed build exceeded a capacity (build again)" );
        return st;
    }

    sesdf_status sesdf_cb_get_stream( sesdf_cb_handle * hh, void ** stream )
    {
        if ( !hh || !stream )
            return report_error( SESDF_ERR_INVALID, "null argument" );
        *stream = hh->h.stream;
        return SESDF_OK;
    }

    sesdf_status sesdf_cb_get_data( sesdf_cb_handle * hh, sesdf_cb_data * out )
    {
        if ( !hh || !out )
            return report_error( SESDF_ERR_INVALID, "null argument" );
        const cb::Handle & h = hh->h;
        if ( !h.built && h.n )
            return report_error( SESDF_ERR_NOT_BUILT, "no build yet" );
        std::memset( out, 0, sizeof( *out ) );
        out->atom_nb                  = h.n;
        out->atoms                    = reinterpret_cast<float *>( h.atoms );
        out->probe_nb                 = h.hctl->probeNb;
        out->probe_positions          = reinterpret_cast<float *>( h.probePos );
        out->sphere_triangle_vec1     = reinterpret_cast<float *>( h.tria1 );
        out->sphere_triangle_vec2     = reinterpret_cast<float *>( h.tria2 );
        out->sphere_triangle_vec3     = reinterpret_cast<float *>( h.tria3 );
        out->torus_nb                 = h.hctl->torusNb;
        out->torus_positions          = reinterpret_cast<float *>( h.torusPos );
        out->torus_axes               = reinterpret_cast<float *>( h.torusAxis );
        out->torus_visibility_spheres = reinterpret_cast<float *>( h.torusVs );
        out->tex_coords               = h.texCoord;
        out->singularity_texture      = h.singTex;
        out->probe_neighbor_nb        = cb::ProbeK;
        out->tex_size                 = h.prm.texSize;
        out->probe_radius             = h.prm.rp;
        return SESDF_OK;
    }

    sesdf_status sesdf_cb_get_counts( sesdf_cb_handle * hh, sesdf_cb_counts * out )
    {
        if ( !hh || !out )
            return report_error( SESDF_ERR_INVALID, "null argument" );
        const cb::Handle & h = hh->h;
        std::memset( out, 0, sizeof( *out ) );
        out->atom_nb                 = h.n;
        out->probe_nb                = h.hctl->probeNb;
        out->torus_nb                = h.hctl->torusNb;
        out->neighbor_overflow_atoms = h.hctl->nbOverflow;
        out->no_contour_atoms        = h.hctl->noContour;
        out->arc_capacity            = h.arcCap;
        out->build_attempts          = h.attempts;
        out->grid_dimensions         = uint32_t( h.prm.dim );
        out->singularity_grid        = h.fineOn ? 1u : 0u;
        out->device_bytes            = h.fixedBytes + h.outBytes;
        out->build_serial            = h.serial;
        return SESDF_OK;
    }

    sesdf_status sesdf_cb_set_tuning( sesdf_cb_handle * hh, const sesdf_cb_tuning * t )
    {
        if ( !hh || !t )
            return report_error( SESDF_ERR_INVALID, "null argument" );
        if ( t->arc_capacity != 0 && t->arc_capacity != cb::ArcCapSmall && t->arc_capacity != cb::ArcCapFull )
            return report_error( SESDF_ERR_INVALID, "arc_capacity must be 0, 8 or 80" );
        if ( t->singularity_search > 2 )
            return report_error( SESDF_ERR_INVALID, "singularity_search must be 0, 1 or 2" );
        cb::Handle & h = hh->h;
        CB_TRY( hipSetDevice( h.device ) );
        CB_TRY( hipStreamSynchronize( h.stream ) );
        h.singSearch      = int( t->singularity_search );
        const bool fineOn = h.singSearch == 2 || ( h.singSearch == 0 && h.fineBig );
        if ( fineOn != h.fineOn )
        {
            h.fineOn = fineOn;
            if ( h.graph )
            {
                CB_TRY( hipGraphExecDestroy( h.graph ) );
                h.graph = nullptr;
            }
        }
        if ( t->arc_capacity && t->arc_capacity != h.arcCap )
        {
            h.arcCap = t->arc_capacity;
            if ( h.graph )
            {
                CB_TRY( hipGraphExecDestroy( h.graph ) );
                h.graph = nullptr;
            }
        }
        if ( t->probe_capacity || t->torus_capacity )
        {
            if ( t->probe_capacity )
                h.capP = t->probe_capacity;
            if ( t->torus_capacity )
                h.capT = t->torus_capacity;
            h.built = false; // the outputs of the last build are gone
            return cb::allocate_outputs( h );
        }
        return SESDF_OK;
    }

    sesdf_status sesdf_cb_get_internal( sesdf_cb_handle * hh, sesdf_cb_internal * out )
    {
        if ( !hh || !out )
            return report_error( SESDF_ERR_INVALID, "null argument" );
        const cb::Handle & h = hh->h;
        out->sorted_to_initial = h.sortedToInitial;
        out->neighbor_count    = h.nbCount;
        out->neighbor_ids      = h.nbIds;
        out->max_neighbors     = cb::K;
        out->probe_atoms       = reinterpret_cast<uint32_t *>( h.probeIds );
        out->torus_atoms       = reinterpret_cast<uint32_t *>( h.torusIds );
        out->probe_sorted      = h.probeSorted;
        return SESDF_OK;
    }
}
