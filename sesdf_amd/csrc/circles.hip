// Circle stage: neighbour query + circle visibility (one wave per atom), then convex sectors,
// intersected-circle ids and full circles from per-atom offsets.
//
// Replaces, for one atom i:
//   findCirclesBetweenAtoms        (sesdf/src/sesdf/operations.cu:34-124; ssesdf/src/ssesdf/operations.cu:77-167)
//   computeCirclesVisibilityStatus (sesdf operations.cu:126-204; ssesdf operations.cu:169-244)
//   the scans / copy_if of findCircles (sesdf/include/bcs/sesdf/operations.inl:32-93)
//   buildConvexPatchesSectors      (operations.inl:175-228)
// Atom i's neighbour atoms are staged in LDS in i's local frame (coordinates relative to c_i, so
// FP32 keeps ~1e-6 A resolution even at |c| ~ 400 A), the candidate rows of the grid stencil are
// flattened across the 64 lanes, and neighbours are appended with a ballot in ascending sorted
// index (the reference's order). Output offsets come from a scan of per-atom counts, so outputs
// are deterministic and no N x K mask scan is needed.
#include <algorithm>

#include "kernels.h"

namespace sesdf
{
    constexpr int CirclesWaves = 4;

    // Visibility predicates of circle (i, j) against sphere k, without sqrt or division.
    // Reference (operations.cu:172-196): with v = centre - c_k, s = v.n, h = |v x n| the in-plane
    // distance, R the circle radius:
    //   hidden  iff (h + R)^2 + s^2 - Rk^2 < eps   <=>  2hR < A
    //   touched iff (R - h)^2 + s^2 - Rk^2 < eps   <=>  2hR > -A
    // with A = Rk^2 + eps - |v|^2 - R^2; both sides are compared squared.
    struct CircleTest
    {
        f3    centre;   // local frame (relative to c_i)
        f3    normal;
        float radius2;
        float fourR2;
    };

    __device__ __forceinline__ CircleTest make_circle( f3 dj, float Ri2, float Rj2 )
    {
        CircleTest  c;
        const float d2 = len2( dj );
        const float t  = ( Ri2 - Rj2 + d2 ) / ( 2.f * d2 );
        c.centre       = dj * t;
        c.radius2      = Ri2 - len2( c.centre );
        c.normal       = dj * ( 1.0f / sqrtf( d2 ) );
        c.fourR2       = 4.f * c.radius2;
        return c;
    }

    // make_circle with hardware reciprocal square root (about 1 ulp): for the visibility
    // decisions, whose epsilon is far wider.
    __device__ __forceinline__ CircleTest make_circle_fast( f3 dj, float Ri2, float Rj2 )
    {
        CircleTest  c;
        const float d2 = len2( dj );
        const float r  = frsq( d2 );
        c.centre       = dj * ( ( Ri2 - Rj2 + d2 ) * ( 0.5f * r * r ) );
        c.radius2      = Ri2 - len2( c.centre );
        c.normal       = dj * r;
        c.fourR2       = 4.f * c.radius2;
        return c;
    }

    // The SAS spheres of i and j meet in no circle when R^2 <= 0: one contains the other, or they
    // are tangent or apart (pairs within the neighbour test's slack of R_i + R_j). The pair then has
    // no circle: status 0. (The reference takes a square root of R^2 < 0 there, and its NaN radius
    // passes no test: status 2, a full circle with NaN geometry.)
    __device__ __forceinline__ bool no_circle( const CircleTest & c ) { return !( c.radius2 > 0.f ); }

    // orthogonalVector (sesdf/include/bcs/cuda/math.cuh:32-38)
    __device__ __forceinline__ f3 orthogonal_vector( f3 n )
    {
        f3 ref = mk( 1.f, 0.f, 0.f );
        if ( fabsf( n.x ) > 1.f - 1e-4f )
            ref = mk( 0.f, 1.f, 0.f );
        return unit( cross( n, ref ) );
    }

    // Circle (i, slot l) tested again against every other sphere of N(i) with h^2 = |c_k x n|^2, for
    // circles with a sphere centre within 0.03 A of their axis. |c_k|^2 - e^2 cancels there: its
    // rounding (a few 1e-6 for |c_k| up to 6 A) moves 2 R h, the side the predicates compare
    // against A, by about R 4e-6 / h, which reaches their 1e-3 scale as h shrinks (above 0.03 A it
    // stays under 4e-4).
    __device__ __forceinline__ void circle_exact( const float4 * atom, const float2 * aux, uint32_t ni, uint32_t l, const CircleTest & c,
                                                  float twoRho, bool & hidden, bool & touched, uint32_t ( &wd )[ 4 ] )
    {
        hidden = false, touched = false;
        wd[ 0 ] = wd[ 1 ] = wd[ 2 ] = wd[ 3 ] = 0u;
        for ( uint32_t k = 0; k < ni && !hidden; k++ )
        {
            if ( k == l )
                continue;
            const float4 ak  = atom[ k ];
            const float2 ax  = aux[ k ];
            const float  e   = dot( c.normal, xyz( ak ) );
            const float  h2  = len2( cross( xyz( ak ), c.normal ) );
            const float  A   = fmaf( twoRho, e, ax.y );
            const float  lhs = c.fourR2 * h2;
            const float  A2  = A * A;
            hidden           = A > 0.f && lhs < A2;
            if ( A > 0.f || lhs > A2 )
            {
                touched = true;
                wd[ 0 ] |= k < 32 ? 1u << k : 0u;
                wd[ 1 ] |= k >= 32 && k < 64 ? 1u << ( k - 32 ) : 0u;
                wd[ 2 ] |= k >= 64 && k < 96 ? 1u << ( k - 64 ) : 0u;
                wd[ 3 ] |= k >= 96 && k < 128 ? 1u << ( k - 96 ) : 0u;
            }
        }
    }

    template<int KMAX, int MODE, int WAVES>
    __global__ void __launch_bounds__( WAVES * WaveSize ) circles_count_kernel( Buffers b )
    {
        __shared__ uint32_t sIdx[ WAVES ][ KMAX ];
        __shared__ float4   sAtom[ WAVES ][ KMAX ];   // (c_k - c_i, Rk^2)
        __shared__ float2   sAux[ WAVES ][ KMAX ];    // (|c_k - c_i|^2, Rk^2 + eps - |c_k - c_i|^2 - Ri^2)
        __shared__ uint32_t sRowStart[ WAVES ][ WaveSize ];
        __shared__ uint32_t sRowPrefix[ WAVES ][ WaveSize ];     // first element of each row
        __shared__ uint32_t sRowOf[ WAVES ][ WaveSize ];         // row marks of one 64-element window

        const int      w    = wave_in_block();
        const int      lane = lane_id();
        const uint32_t i    = xcd_block() * WAVES + w;
        if ( i >= b.n )
            return;
        const float    rp   = b.probeRadius;
        const float    rmax = __uint_as_float( b.ctl->radiusMaxBits );
        const uint32_t K    = b.K;

        const float4   ai  = b.atoms[ i ];
        const f3       ci  = xyz( ai );
        const float    Ri  = ai.w + rp;
        const float    Ri2 = Ri * Ri;
        const Molecule M   = b.mols[ b.molNb > 1 ? b.atomMol[ i ] : 0u ];

        // ---- neighbour query over the flattened rows of the cell stencil -------------------------
        const float range = ai.w + rmax + 2.f * rp + 1e-3f;
        const int   reach = max( int( ceilf( range / M.cell ) ), 1 );
        int         gx, gy, gz;
        cell_of( M, ci.x, ci.y, ci.z, gx, gy, gz );
        const int x0 = max( gx - reach, 0 ), x1 = min( gx + reach, M.dx - 1 );
        const int y0 = max( gy - reach, 0 ), y1 = min( gy + reach, M.dy - 1 );
        const int z0 = max( gz - reach, 0 ), z1 = min( gz + reach, M.dz - 1 );
        const int ny = y1 - y0 + 1, rows = ny * ( z1 - z0 + 1 );

        uint32_t count = 0;
        for ( int rowBase = 0; rowBase < rows; rowBase += WaveSize )
        {
            const int r  = rowBase + lane;
            uint32_t  st = 0, len = 0;
            if ( r < rows )
            {
                // the row's cells within `range` of c_i: its y/z gap leaves a half-width along x
                const int   yy  = y0 + r % ny, zz = z0 + r / ny;
                const float gyv = slab_gap( M.oy + M.cell * float( yy ), M.cell, ci.y );
                const float gzv = slab_gap( M.oz + M.cell * float( zz ), M.cell, ci.z );
                const float rem = range * range - gyv * gyv - gzv * gzv;
                if ( rem >= 0.f )
                {
                    const float half = sqrtf( rem );
                    const int   xa   = max( x0, int( floorf( ( ci.x - half - M.ox ) / M.cell ) ) );
                    const int   xb   = min( x1, int( floorf( ( ci.x + half - M.ox ) / M.cell ) ) );
                    if ( xa <= xb )
                    {
                        const uint32_t h0 = cell_index( M, xa, yy, zz );
                        st                = b.cellStart[ h0 ];
                        len               = b.cellStart[ h0 + uint32_t( xb - xa ) + 1 ] - st;
                    }
                }
            }
            const uint32_t incl  = wave_incl_scan( len );
            const uint32_t excl  = incl - len;
            const uint32_t total = wave_bcast( incl, WaveSize - 1 );
            sRowStart[ w ][ lane ]  = st;
            sRowPrefix[ w ][ lane ] = excl;
            for ( uint32_t base = 0; base < total; base += WaveSize )
            {
                // Row of element e = base + lane: each non-empty row of the window [base, base + 64)
                // marks its first element there with its index (the row holding element `base`
                // marks slot 0), and a prefix maximum carries the marks forward, since rows
                // ascend with their elements.
                sRowOf[ w ][ lane ] = 0u;
                wave_sync();
                if ( len > 0u && excl < base + WaveSize && incl > base )
                    sRowOf[ w ][ max( excl, base ) - base ] = uint32_t( lane );
                wave_sync();
                const int      row  = wave_incl_max( int( sRowOf[ w ][ lane ] ) );
                const uint32_t e    = base + lane;
                bool           pass = false;
                uint32_t       j    = 0;
                float4         aj   = make_float4( 0.f, 0.f, 0.f, 0.f );
                if ( e < total )
                {
                    j                = sRowStart[ w ][ row ] + ( e - sRowPrefix[ w ][ row ] );
                    aj               = b.atoms[ j ];
                    const f3    d    = xyz( aj ) - ci;
                    const float Rs   = Ri + ( aj.w + rp );
                    pass             = j != i && len2( d ) - Rs * Rs <= PredicateEps;
                    aj.x = d.x, aj.y = d.y, aj.z = d.z;
                    aj.w = ( aj.w + rp ) * ( aj.w + rp );
                }
                const uint64_t mask = ballot( pass );
                const uint32_t slot = count + rank_below( mask );
                if ( pass && slot < K )
                {
                    sIdx[ w ][ slot ]  = j;
                    sAtom[ w ][ slot ] = aj;
                }
                count += uint32_t( __popcll( mask ) );
            }
            wave_sync();
        }
        if ( count > K && lane == 0 )
        {
            atomicAdd( &b.ctl->nbOverflowAtoms, 1u ); // rare: truncated list, as the reference
            atomicMax( &b.ctl->nbMaxCount, count );
        }
        const uint32_t ni = uint32_t( __builtin_amdgcn_readfirstlane( int( min( count, K ) ) ) ); // uniform
        wave_sync();

        if ( lane == 0 )
            b.nbCount[ i ] = ni;
        for ( uint32_t l = lane; l < ni; l += WaveSize )
        {
            b.nbIds[ size_t( i ) * K + l ] = sIdx[ w ][ l ];
            const float4 ak  = sAtom[ w ][ l ];
            const float  K2  = len2( xyz( ak ) );
            sAux[ w ][ l ]   = make_float2( K2, ( ak.w + PredicateEps ) - K2 - Ri2 );
        }
        // up to 8 padding slots past ni that neither hide, touch nor come near any circle: e = 0,
        // h^2 = 1e30, A = -1e30 (so A |A| = -inf)
        if ( uint32_t( lane ) < min( 8u, uint32_t( KMAX ) - ni ) )
        {
            sAtom[ w ][ ni + lane ] = make_float4( 0.f, 0.f, 0.f, 0.f );
            sAux[ w ][ ni + lane ]  = make_float2( 1e30f, -1e30f );
        }
        wave_sync();

        // ---- visibility of circles (i, j): lanes over j, sequential over k in LDS -----------------
        // The circle centre lies on the axis n (centre = rho n) and its points on sphere i
        // (rho^2 + R^2 = Ri^2), so with e = n . c_k (c_k relative to c_i):
        //   h^2 = |(centre - c_k) x n|^2 = |c_k|^2 - e^2,
        //   A   = Rk^2 + eps - |centre - c_k|^2 - R^2 = (Rk^2 + eps - |c_k|^2 - Ri^2) + 2 rho e,
        // one dot product and three fused operations per (circle, sphere) instead of a cross product
        // and two squared lengths.
        uint32_t nV = 0, nT = 0, nF = 0;
        // The exterior path tests only circles with j > i: the suffix of the ascending list from s0.
        uint32_t s0 = 0;
        if ( MODE == 1 )
            for ( uint32_t lb = 0; lb < ni; lb += WaveSize )
                s0 += uint32_t( __popcll( ballot( lb + lane < ni && sIdx[ w ][ lb + lane ] < i ) ) );
        const uint32_t nc = ni - s0; // circles to test
        if ( nc > 0 && nc <= 32 && ni <= 64 )
        {
            // Up to 32 circles: both half-waves take the same circle (lane & 31) and split the
            // spheres, the lower half k in [0, hk), the upper [hk, 2 hk) (inert padding past ni),
            // so the lockstep loop is half as long. Hidden / near-axis states join across the
            // halves in the wave masks, the touch words by one lane exchange.
            const uint32_t   c      = uint32_t( lane ) & 31u;
            const bool       upper  = lane >= 32;
            const uint32_t   l      = s0 + c;
            const bool       act    = c < nc;
            const uint32_t   j      = act ? sIdx[ w ][ l ] : 0u;
            const float4     aj     = sAtom[ w ][ act ? l : 0u ];
            const CircleTest cg     = make_circle_fast( xyz( aj ), Ri2, aj.w );
            const float      twoRho = 2.f * dot( cg.centre, cg.normal );
            const float      fourR2 = fmaxf( cg.fourR2, 0.f );
            const uint32_t   hk     = ( ( ni + 1u ) / 2u + 3u ) & ~3u; // <= 32; 2 hk <= ni + 7
            const uint32_t   off    = upper ? hk : 0u;
            const uint64_t   actM   = ballot( act ) & 0xffffffffull;
            uint64_t         hidM = 0, nearM = 0;
            uint32_t         tw   = 0; // touch bits of slots off + [0, hk)
            for ( uint32_t kk = 0; kk < hk; kk += 4 )
            {
#pragma unroll
                for ( uint32_t u = 0; u < 4; u++ )
                {
                    const uint32_t k     = kk + u + off;
                    const float4   ak    = sAtom[ w ][ k ];
                    const float2   ax    = sAux[ w ][ k ];
                    const float    e     = fmaf( cg.normal.z, ak.z, fmaf( cg.normal.y, ak.y, cg.normal.x * ak.x ) );
                    const float    h2    = fmaf( -e, e, ax.x );
                    const float    A     = fmaf( twoRho, e, ax.y );
                    const float    lhs   = fourR2 * h2;
                    const float    P     = A * fabsf( A );
                    const bool     other = k != l;
                    nearM |= ballot( other && h2 < 1e-3f );
                    hidM |= ballot( other && lhs < P );
                    tw |= other && lhs > -P ? 1u << ( kk + u ) : 0u;
                }
                if ( ( actM & ~( hidM | ( hidM >> 32 ) ) ) == 0 )
                    break; // every circle is hidden: the rest no longer matters
            }
            const uint64_t hidC  = hidM | ( hidM >> 32 );
            const uint64_t nearC = nearM | ( nearM >> 32 );
            bool           hidden  = ( hidC >> c ) & 1u;
            const bool     near    = ( nearC >> c ) & 1u;
            const bool     none    = no_circle( cg );
            const uint32_t twUpper = __shfl_xor( tw, 32, WaveSize );
            const uint64_t T       = upper ? 0ull : ( uint64_t( tw ) | ( uint64_t( twUpper ) << hk ) ); // slots [0, 2 hk)
            uint32_t       wd[ 4 ] = { uint32_t( T ), uint32_t( T >> 32 ), 0u, 0u };
            bool           touched = T != 0ull;
            if ( __builtin_expect( ballot( !upper && act && near ) != 0, 0 ) )
                if ( !upper && act && near )
                    circle_exact( sAtom[ w ], sAux[ w ], ni, l, cg, twoRho, hidden, touched, wd );
            const bool     own = !upper && act;
            hidden             = hidden || none;
            const uint32_t st  = own ? ( hidden ? 0u : ( touched ? 1u : 2u ) ) : 0u;
            for ( uint32_t z = lane; z < s0; z += WaveSize ) // exterior: circles with j < i are not evaluated
                b.status[ size_t( i ) * K + z ] = 0u;
            if ( own )
                b.status[ size_t( i ) * K + l ] = uint8_t( st );
            const uint64_t mT = ballot( own && st == 1 && j > i );
            if ( own && st == 1 && j > i )
            {
                const uint32_t r = rank_below( mT );
                if ( r < TouchStride && K <= 128 )
                    b.touch[ size_t( i ) * TouchStride + r ] = make_uint4( wd[ 0 ], wd[ 1 ], wd[ 2 ], wd[ 3 ] );
            }
            nV = uint32_t( __popcll( ballot( own && st != 0 ) ) );
            nT = uint32_t( __popcll( mT ) );
            nF = uint32_t( __popcll( ballot( own && st == 2 && j > i ) ) );
            if ( MODE == 0 && lane == 0 )
                b.intMask[ size_t( i ) * ( K / 32 ) ] = uint32_t( mT ); // s0 = 0: lane c is slot c
        }
        // slots past ni are never read (statuses, touch sets, intersected-circle bits: all readers
        // stop at the neighbour count)
        else for ( uint32_t lb = 0; lb < ni; lb += WaveSize )
        {
            const uint32_t l      = lb + lane;
            const uint32_t j      = l < ni ? sIdx[ w ][ l ] : 0u;
            const bool     active = l < ni && ( MODE == 0 || i < j );
            uint32_t       st     = 0;
            uint64_t       touch0 = 0, touch1 = 0; // spheres touching circle (i, j), by slot (K <= 128)
            if ( __builtin_expect( ballot( active ) != 0, 1 ) )
            {
                const float4     aj     = sAtom[ w ][ min( l, K - 1 ) ];
                const CircleTest c      = make_circle_fast( xyz( aj ), Ri2, aj.w );
                const float      twoRho = 2.f * dot( c.centre, c.normal );
                const float      fourR2 = fmaxf( c.fourR2, 0.f );
                // Lanes run the k loop in lockstep (k is wave-uniform). With P = A |A| and lhs >= 0
                // (h^2 >= 0 wherever the test below is not redone):
                //   hidden  <=> A > 0 and lhs < A^2   <=> lhs <  P
                //   touched <=> A > 0 or  lhs > A^2   <=> lhs > -P
                // Hidden / near-axis states are wave masks (scalar); a lane's own slot and the
                // padding slots past ni are masked out of them, and out of the touch words at the
                // end of each 32-slot block. The wave leaves once every active lane is hidden.
                const uint64_t activeM = ballot( active );
                uint64_t       hidM = 0, nearM = 0;
                uint32_t       w0 = 0, w1 = 0, w2 = 0, w3 = 0, anyT = 0, word = 0;
                const uint32_t nPad = ( ni + 3u ) & ~3u; // <= KMAX (a multiple of 4); padding is inert
                auto           test = [ & ]( uint32_t k, uint64_t self ) {
                    const float4 ak  = sAtom[ w ][ k ];
                    const float2 ax  = sAux[ w ][ k ];
                    const float  e   = fmaf( c.normal.z, ak.z, fmaf( c.normal.y, ak.y, c.normal.x * ak.x ) );
                    const float  h2  = fmaf( -e, e, ax.x );
                    const float  A   = fmaf( twoRho, e, ax.y );
                    const float  lhs = fourR2 * h2;
                    const float  P   = A * fabsf( A );
                    nearM |= ballot( h2 < 1e-3f ) & ~self;
                    hidM |= ballot( lhs < P ) & ~self;
                    word |= lhs > -P ? 1u << ( k & 31u ) : 0u;
                };
                for ( uint32_t k0 = 0; k0 < nPad; k0 += 4 )
                {
                    if ( k0 - lb < uint32_t( WaveSize ) ) // slots of this pass's lanes: skip each own
                    {
#pragma unroll
                        for ( uint32_t u = 0; u < 4; u++ )
                            test( k0 + u, uint64_t( 1 ) << ( k0 + u - lb ) );
                    }
                    else
                    {
#pragma unroll
                        for ( uint32_t u = 0; u < 4; u++ )
                            test( k0 + u, 0ull );
                    }
                    if ( ( k0 & 31u ) == 28u || k0 + 4 >= nPad )
                    {
                        // block q of slots [32q, 32q + 32): keep real slots other than the lane's own
                        const uint32_t q     = k0 >> 5;
                        const uint32_t valid = ni - 32u * q >= 32u ? ~0u : ( 1u << ( ni - 32u * q ) ) - 1u;
                        word &= valid & ~( ( l >> 5 ) == q ? 1u << ( l & 31u ) : 0u );
                        w0 = q == 0 ? word : w0;
                        w1 = q == 1 ? word : w1;
                        w2 = q == 2 ? word : w2;
                        w3 = q == 3 ? word : w3;
                        anyT |= word;
                        word = 0u;
                    }
                    if ( ( activeM & ~hidM ) == 0 )
                        break; // every active circle is hidden: the rest no longer matters
                }
                bool hidden  = ( hidM >> lane ) & 1u;
                bool near    = ( nearM >> lane ) & 1u;
                bool touched = anyT != 0u;
                if ( __builtin_expect( ballot( active && near ) != 0, 0 ) )
                {
                    if ( active && near )
                    {
                        uint32_t wd[ 4 ];
                        circle_exact( sAtom[ w ], sAux[ w ], ni, l, c, twoRho, hidden, touched, wd );
                        w0 = wd[ 0 ], w1 = wd[ 1 ], w2 = wd[ 2 ], w3 = wd[ 3 ];
                    }
                }
                touch0 = uint64_t( w0 ) | ( uint64_t( w1 ) << 32 );
                touch1 = uint64_t( w2 ) | ( uint64_t( w3 ) << 32 );
                hidden = hidden || no_circle( c );
                st     = active ? ( hidden ? 0u : ( touched ? 1u : 2u ) ) : 0u;
            }
            if ( l < ni )
                b.status[ size_t( i ) * K + l ] = uint8_t( st );
            const uint64_t mT = ballot( l < ni && st == 1 && j > i );
            if ( l < ni && st == 1 && j > i )
            {
                // an occluder of a probe on this circle must touch it: keep the set for the
                // intersection stage (candidate order = slot order of the intersected circles)
                const uint32_t r = nT + rank_below( mT );
                if ( r < TouchStride && K <= 128 )
                    b.touch[ size_t( i ) * TouchStride + r ] =
                        make_uint4( uint32_t( touch0 ), uint32_t( touch0 >> 32 ), uint32_t( touch1 ), uint32_t( touch1 >> 32 ) );
            }
            nV += uint32_t( __popcll( ballot( l < ni && st != 0 ) ) );
            nT += uint32_t( __popcll( mT ) );
            nF += uint32_t( __popcll( ballot( l < ni && st == 2 && j > i ) ) );
            if ( MODE == 0 && lane == 0 )
            {
                b.intMask[ size_t( i ) * ( K / 32 ) + lb / 32 ] = uint32_t( mT );
                if ( lb / 32 + 1 < K / 32 )
                    b.intMask[ size_t( i ) * ( K / 32 ) + lb / 32 + 1 ] = uint32_t( mT >> 32 );
            }
        }
        if ( lane == 0 )
        {
            b.atomCnt[ i ]                     = nV;
            b.atomCnt[ 3 * size_t( b.n ) + i ] = ni; // summed by the scan (neighbour pairs)
            if ( MODE == 0 )
            {
                b.atomCnt[ b.n + i ]     = nT;
                b.atomCnt[ 2 * b.n + i ] = nF;
            }
        }
    }

    // Sectors, intersected circles and full circles of atom i at the offsets of the scan. Four
    // atoms per wave, 16 lanes each (about 45 slots: three rounds); ranks come from the wave's
    // ballots restricted to the atom's 16 lanes.
    constexpr int EmitAtomsPerWave = 4;
    __global__ void __launch_bounds__( CirclesWaves * WaveSize ) circles_emit_kernel( Buffers b )
    {
        const uint32_t lane = uint32_t( lane_id() ), sub = lane & 15u, grp = lane >> 4;
        const uint32_t i    = ( xcd_block() * CirclesWaves + uint32_t( wave_in_block() ) ) * EmitAtomsPerWave + grp;
        const bool     atom = i < b.n;
        const uint64_t mine = uint64_t( 0xffff ) << ( 16 * grp );
        const uint32_t K    = b.K;
        const float    rp   = b.probeRadius;
        const uint32_t ni   = atom ? b.nbCount[ i ] : 0u;
        uint32_t       runV = 0, runT = 0, runF = 0;
        float4         ai   = make_float4( 0.f, 0.f, 0.f, 0.f );
        if ( atom )
        {
            runV = b.vStart[ i ];
            if ( sub == 0 )
                b.convexPatches[ i ] = make_uint2( runV, b.vStart[ i + 1 ] );
            if ( ni )
            {
                runT = b.trimBase[ i ];
                runF = b.fullBase[ i ];
                ai   = b.atoms[ i ];
            }
        }
        const f3       ci     = xyz( ai );
        const float    Ri     = ai.w + rp;
        const float    Ri2    = Ri * Ri;
        uint32_t       most   = ( ni + 15u ) / 16u; // rounds of the wave: its fullest atom's
        for ( int o = 32; o > 0; o >>= 1 )
            most = max( most, uint32_t( __shfl_xor( int( most ), o, WaveSize ) ) );
        const uint32_t rounds = uint32_t( __builtin_amdgcn_readfirstlane( int( most ) ) );
        for ( uint32_t r = 0; r < rounds; r++ )
        {
            const uint32_t l    = 16u * r + sub;
            const bool     live = l < ni;
            const uint32_t st   = live ? b.status[ size_t( i ) * K + l ] : 0u;
            const uint32_t j    = live ? b.nbIds[ size_t( i ) * K + l ] : 0u;
            const bool     isV = live && st != 0, isT = live && st == 1 && j > i, isF = live && st == 2 && j > i;
            const uint64_t mV = ballot( isV ) & mine, mT = ballot( isT ) & mine, mF = ballot( isF ) & mine;
            if ( isV )
            {
                const uint32_t dst = runV + rank_below( mV );
                if ( dst < b.capSectors )
                {
                    const float4     aj = b.atoms[ j ];
                    const float      Rj = aj.w + rp;
                    // half angle of the cap: acos(n . x / Ri) for a point x of the circle
                    // (buildConvexPatchesSectors, operations.inl:175-228, takes x = centre +
                    // radius * orthogonalVector(n)); n . x is the plane offset n . centre. For
                    // spheres tangent within the predicate's epsilon (radius^2 < 0) the
                    // reference's radius is NaN and its clamp (fmaxf(-1, fminf(v, 1))) makes the
                    // angle 0; the plane offset is then Ri or just above it: the same 0.
                    const CircleTest c  = make_circle( xyz( aj ) - ci, Ri2, Rj * Rj );
                    const float      ca = fmaxf( -1.f, fminf( dot( c.normal, c.centre ) / Ri, 1.f ) );
                    b.sectors[ dst ]    = make_float4( c.normal.x, c.normal.y, c.normal.z, acosf( ca ) );
                }
                else
                    set_overflow( b.ctl, OverflowSectors );
            }
            if ( isT )
            {
                const uint32_t dst = runT + rank_below( mT );
                if ( dst < b.capTrimmed )
                {
                    b.trimmedToGlobal[ dst ] = i * K + l;
                    b.circleIntNb[ dst ]     = 0u;
                }
                else
                    set_overflow( b.ctl, OverflowTrimmed );
            }
            if ( isF )
            {
                const uint32_t dst = runF + rank_below( mF );
                if ( dst < b.capFull )
                    b.fullCircles[ dst ] = make_uint2( i, j );
                else
                    set_overflow( b.ctl, OverflowFullCircles );
            }
            runV += uint32_t( __popcll( mV ) );
            runT += uint32_t( __popcll( mT ) );
            runF += uint32_t( __popcll( mF ) );
        }
    }

    int device_cu_count()
    {
        static int cached = 0;
        if ( cached )
            return cached;
        int dev = 0, cus = 0;
        if ( hipGetDevice( &dev ) != hipSuccess )
            return 256;
        if ( hipDeviceGetAttribute( &cus, hipDeviceAttributeMultiprocessorCount, dev ) != hipSuccess || cus <= 0 )
            return 256;
        cached = cus;
        return cus;
    }

    template<int KMAX, int MODE, int WAVES>
    static void launch_count_shape( const Buffers & b, hipStream_t s )
    {
        circles_count_kernel<KMAX, MODE, WAVES><<<( b.n + WAVES - 1 ) / WAVES, WAVES * WaveSize, 0, s>>>( b );
    }

    template<int KMAX, int MODE>
    static void launch_count_waves( const Buffers & b, hipStream_t s )
    {
        // block shape (waves per workgroup) chosen by the autotuner
        if ( b.wavesCircles == 2 )
            launch_count_shape<KMAX, MODE, 2>( b, s );
        else if ( b.wavesCircles == 8 )
            launch_count_shape<KMAX, MODE, 8>( b, s );
        else
            launch_count_shape<KMAX, MODE, 4>( b, s );
    }

    static void launch_circles_count( const Buffers & b, hipStream_t s )
    {
        if ( b.K <= 128 )
        {
            if ( b.mode == 0 )
                launch_count_waves<128, 0>( b, s );
            else
                launch_count_waves<128, 1>( b, s );
        }
        else
        {
            if ( b.mode == 0 )
                launch_count_waves<320, 0>( b, s );
            else
                launch_count_waves<320, 1>( b, s );
        }
    }

    void launch_circles( const Buffers & b, hipStream_t s )
    {
        if ( b.n == 0 )
            return;
        launch_circles_count( b, s );
    }

    void launch_circles_finish( const Buffers & b, hipStream_t s )
    {
        if ( b.n == 0 )
            return;
        const uint32_t perBlock = CirclesWaves * EmitAtomsPerWave;
        const uint32_t blocks   = ( b.n + perBlock - 1 ) / perBlock;
        ScanArgs       a {};
        a.n           = b.n;
        a.ticket      = TicketCircles;
        a.stage       = 2;
        a.in[ 0 ]     = b.atomCnt;
        a.out[ 0 ]    = b.vStart;
        a.totals[ 0 ] = b.mode == 0 ? &b.ctl->sectorNb : &b.ctl->visibleCircleNb;
        a.sumOut      = &b.ctl->neighborNb;
        if ( b.mode == 0 )
        {
            a.in[ 1 ]     = b.atomCnt + b.n;
            a.in[ 2 ]     = b.atomCnt + 2 * size_t( b.n );
            a.in[ 3 ]     = b.atomCnt + 3 * size_t( b.n ); // neighbour counts (summed)
            a.out[ 1 ]    = b.trimBase;
            a.out[ 2 ]    = b.fullBase;
            a.totals[ 1 ] = &b.ctl->intersectedCircleNb;
            a.totals[ 2 ] = &b.ctl->fullCircleNb;
            launch_scan<3>( b, a, s );
            circles_emit_kernel<<<blocks, CirclesWaves * WaveSize, 0, s>>>( b );
        }
        else
        {
            a.in[ 1 ] = b.atomCnt + 3 * size_t( b.n ); // neighbour counts (summed)
            launch_scan<1>( b, a, s );
        }
    }

    // The reference's globalToTrimmedId (sesdf/include/bcs/sesdf/operations.inl:46-70): for every
    // neighbour slot g = i * K + l of the N * K + 1 slots, the exclusive prefix count of the
    // intersected circles (status 1, j > i) before it. trimmedToGlobal lists those circles' slots in
    // ascending order, so the count is a lower bound in it.
    __global__ void global_to_trimmed_kernel( Buffers b, uint32_t * out, uint64_t slots )
    {
        const uint32_t Ci = min( b.ctl->intersectedCircleNb, b.capTrimmed );
        for ( uint64_t g = uint64_t( blockIdx.x ) * blockDim.x + threadIdx.x; g < slots; g += uint64_t( gridDim.x ) * blockDim.x )
        {
            uint32_t lo = 0, hi = Ci;
            while ( lo < hi )
            {
                const uint32_t mid = ( lo + hi ) >> 1;
                if ( uint64_t( b.trimmedToGlobal[ mid ] ) < g )
                    lo = mid + 1;
                else
                    hi = mid;
            }
            out[ g ] = lo;
        }
    }

    void launch_global_to_trimmed( const Buffers & b, uint32_t * out, hipStream_t s )
    {
        const uint64_t slots  = uint64_t( b.n ) * b.K + 1;
        const uint32_t blocks = uint32_t( std::min<uint64_t>( ( slots + 255 ) / 256, uint64_t( device_cu_count() ) * 16 ) );
        global_to_trimmed_kernel<<<blocks, 256, 0, s>>>( b, out, slots );
    }
} // namespace sesdf
