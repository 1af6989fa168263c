// Concave patches, toroidal segments, probe neighbourhoods and exterior torus records.
//
//   intersections_count_kernel <- findIntersectionsBetweenCircles (sesdf/include/bcs/sesdf/operations.inl:230-486;
//                            ssesdf/include/bcs/ssesdf/operation.inl:100-351): one wave per atom i,
//                            candidate pairs (ij, ik) flattened over the lanes, k in N(j) by the
//                            predicate that built j's list (binary search only after a truncation),
//                            occluders N(i) from LDS; records staged per atom (4 bytes each).
//   intersections_emit_kernel <- the |w| sort of findIntersections (operations.inl:128-141), the
//                            slot of k in j's list and the per-circle counts; offsets from a scan of
//                            per-atom counts.
//   launch_circle_scan    <- the per-circle offset scans (operations.inl:149-156, 500-517)
//   fill_kernel           <- fillIntersections (sesdf/src/sesdf/operations.cu:206-269;
//                            ssesdf/src/ssesdf/operations.cu:287-329)
//   segments_kernel       <- buildSegments (operations.inl:552-668)
//   probe_neighbors_kernel<- handleIntersectionSingularities + findIntersectionsNeighbors
//                            (sesdf operations.cu:14-32, 271-338; ssesdf operations.cu:57-75, 331-399)
//   torus_kernel          <- fillCircles + its count/copy_if (ssesdf operations.cu:246-285,
//                            operation.inl:73-95)
#include "kernels.h"

namespace sesdf
{
    constexpr uint32_t RecMinus    = 1u << 31;
    constexpr uint32_t RecHasNb    = 1u << 30;
    constexpr uint32_t PairRound   = 192; // candidate pairs per compaction round (LDS list per wave;
                                          // 192 keeps a 4-wave block under 22.8 KB: 7 blocks per CU)

    struct CircleGeo
    {
        f3    centre;
        f3    normal;
        float radius2;
    };
    __device__ __forceinline__ CircleGeo circle_geo( f3 dj, float Ri2, float Rj2 )
    {
        CircleGeo   c;
        const float d2 = len2( dj );
        const float t  = ( Ri2 - Rj2 + d2 ) / ( 2.f * d2 );
        c.centre       = dj * t;
        c.radius2      = Ri2 - len2( c.centre );
        c.normal       = dj * ( 1.0f / sqrtf( d2 ) );
        return c;
    }

    // circle_geo with hardware reciprocal square root (about 1 ulp), for decisions only.
    __device__ __forceinline__ CircleGeo circle_geo_fast( f3 dj, float Ri2, float Rj2 )
    {
        CircleGeo   c;
        const float d2 = len2( dj );
        const float r  = frsq( d2 );
        c.normal       = dj * r;
        c.centre       = dj * ( ( Ri2 - Rj2 + d2 ) * ( 0.5f * r * r ) );
        c.radius2      = Ri2 - len2( c.centre );
        return c;
    }

    // Ill-conditioned triplets (near-collinear centres or discriminant near zero) are re-solved in
    // FP64: the circle-circle intersection point xc and discriminant t of the probe trilateration.
    __device__ __forceinline__ bool ill_conditioned( float den, float t, float Ri2 ) { return den < 1e-3f || fabsf( t ) < 1e-3f * Ri2; }
    // Probe positions (the fill) take a stricter flag: an FP32 solve errs by about its rounding
    // (~1e-6 A in i's frame) times 1 / sqrt(den), and the discriminant's root by its rounding over
    // 2 sqrt(t); below these thresholds that would approach the 1e-4 A the positions are held to.
    __device__ __forceinline__ bool ill_conditioned_position( float den, float t, float Ri2 ) { return den < 3e-2f || fabsf( t ) < 3e-2f * Ri2; }

    // c_b - c_a in FP64 from the stored coordinates: exact, where the FP32 difference rounds (its
    // operands differ by more than a factor of two) and an ill-conditioned solve would amplify that.
    __device__ __forceinline__ d3 rel64( float4 b, float4 a ) { return mkd( xyz( b ) ) - mkd( xyz( a ) ); }

    struct Resolved
    {
        f3   x0, x1;
        bool ok;
    };

    // FP64 re-solve of xc, t and the two probe positions of triplet (i, j, k), in i's local frame,
    // from the stored (absolute) coordinates. Out of line: its double temporaries would otherwise
    // set the register budget, and so the occupancy, of the whole pair loop it is rarely taken in.
    __device__ __noinline__ Resolved resolve_fp64( float4 ai, float4 aj, float4 ak, float rp )
    {
        Resolved     r {};
        const double RiD = double( ai.w ) + double( rp );
        const double RjD = double( aj.w ) + double( rp );
        const double RkD = double( ak.w ) + double( rp );
        const d3     Dj = rel64( aj, ai ), Dk = rel64( ak, ai );
        const double dj2 = dot( Dj, Dj ), dk2 = dot( Dk, Dk );
        const d3     cij = Dj * ( ( RiD * RiD - RjD * RjD + dj2 ) / ( 2.0 * dj2 ) );
        const d3     cik = Dk * ( ( RiD * RiD - RkD * RkD + dk2 ) / ( 2.0 * dk2 ) );
        const d3     nij = Dj * ( 1.0 / sqrt( dj2 ) ), nik = Dk * ( 1.0 / sqrt( dk2 ) );
        const d3     uu  = nik - nij * dot( nik, nij );
        const d3     xcD = cij + uu * ( dot( cik - cij, nik ) / dot( uu, nik ) );
        const double tD  = RiD * RiD - dot( xcD, xcD );
        r.ok             = tD >= 0.0;
        if ( r.ok )
        {
            const d3 nn = unit( cross( nik, nij ) ) * sqrt( tD );
            const d3 x1 = xcD - nn, x2 = xcD + nn;
            r.x0        = mk( float( x1.x ), float( x1.y ), float( x1.z ) );
            r.x1        = mk( float( x2.x ), float( x2.y ), float( x2.z ) );
        }
        return r;
    }

    // Probe position of triplet (i, j, k) in i's local frame, mirroring the arithmetic of
    // fillIntersections: t = (c_ik - c_ij).n_ik / (u.n_ik), x1 = c_ij + t u,
    // x = x1 + side sqrt(max(R_i^2 - |x1|^2, 0)) n(n_ij x n_ik). Dj, Dk: c_j - c_i, c_k - c_i.
    __device__ d3 probe_fp64( d3 Dj, d3 Dk, float ri, float rj, float rk, float rp, int side )
    {
        const double Ri = double( ri ) + double( rp ), Rj = double( rj ) + double( rp ), Rk = double( rk ) + double( rp );
        const double dj2 = dot( Dj, Dj ), dk2 = dot( Dk, Dk );
        const d3     cij = Dj * ( ( Ri * Ri - Rj * Rj + dj2 ) / ( 2.0 * dj2 ) );
        const d3     cik = Dk * ( ( Ri * Ri - Rk * Rk + dk2 ) / ( 2.0 * dk2 ) );
        const d3     nij = Dj * ( 1.0 / sqrt( dj2 ) ), nik = Dk * ( 1.0 / sqrt( dk2 ) );
        const d3     u   = nik - nij * dot( nik, nij );
        const double t   = dot( cik - cij, nik ) / dot( u, nik );
        const d3     x1  = cij + u * t;
        const double c   = fmax( Ri * Ri - dot( x1, x1 ), 0.0 );
        const d3     n   = unit( cross( nij, nik ) );
        return x1 + n * ( double( side ) * sqrt( c ) );
    }

    __device__ __forceinline__ f3 probe_fp32( f3 dj, f3 dk, float Ri2, float Rj2, float Rk2, int side, bool & ill )
    {
        const CircleGeo ij = circle_geo( dj, Ri2, Rj2 );
        const CircleGeo ik = circle_geo( dk, Ri2, Rk2 );
        const f3        u  = ik.normal - ij.normal * dot( ik.normal, ij.normal );
        const float     den = dot( u, ik.normal );
        const float     t   = dot( ik.centre - ij.centre, ik.normal ) / den;
        const f3        x1  = ij.centre + u * t;
        const float     c   = Ri2 - len2( x1 );
        ill                 = ill_conditioned_position( den, c, Ri2 );
        const f3 n          = unit( cross( ij.normal, ik.normal ) );
        return x1 + n * ( float( side ) * sqrtf( fmaxf( c, 0.f ) ) );
    }

    __device__ __forceinline__ uint32_t trimmed_id( const Buffers & b, uint32_t atom, uint32_t slot )
    {
        const uint32_t * m   = b.intMask + size_t( atom ) * ( b.K / 32 );
        uint32_t         acc = b.trimBase[ atom ];
        for ( uint32_t w = 0; w < slot / 32; w++ )
            acc += __popc( m[ w ] );
        return acc + __popc( m[ slot / 32 ] & ( ( 1u << ( slot % 32 ) ) - 1u ) );
    }

    __device__ __forceinline__ bool is_intersected( const Buffers & b, uint32_t atom, uint32_t slot )
    {
        return ( b.intMask[ size_t( atom ) * ( b.K / 32 ) + slot / 32 ] >> ( slot % 32 ) ) & 1u;
    }

    __device__ __forceinline__ int find_in_list( const uint32_t * list, uint32_t n, uint32_t key )
    {
        int lo = 0, hi = int( n ) - 1;
        while ( lo <= hi )
        {
            const int      mid = ( lo + hi ) >> 1;
            const uint32_t v   = list[ mid ];
            if ( v == key )
                return mid;
            if ( v < key )
                lo = mid + 1;
            else
                hi = mid - 1;
        }
        return -1;
    }

    template<int KMAX, int WAVES>
    __global__ void __launch_bounds__( WAVES * WaveSize ) intersections_count_kernel( Buffers b )
    {
        __shared__ float4   sAtom[ WAVES ][ KMAX ];   // (c_l - c_i, R_l^2)
        __shared__ float4   sAbs[ WAVES ][ KMAX ];    // (c_l, r_l) as stored
        __shared__ uint16_t sCand[ WAVES ][ KMAX ];
        __shared__ uint4    sTouch[ WAVES ][ TouchStride ];
        __shared__ uint32_t sPair[ WAVES ][ PairRound ];

        const int      w    = wave_in_block();
        const int      lane = lane_id();
        const uint32_t i    = xcd_block() * WAVES + w;
        if ( i >= b.n )
            return;
        const float    rp   = b.probeRadius;
        const uint32_t K    = b.K;
        const float4   ai   = b.atoms[ i ];
        const f3       ci   = xyz( ai );
        const float    Ri   = ai.w + rp;
        const float    Ri2  = Ri * Ri;
        const uint32_t ni   = b.nbCount[ i ];
        uint32_t * const stage = b.stage + size_t( i ) * b.stageStride;
        const bool       truncated = b.ctl->nbOverflowAtoms != 0;

        uint32_t m = 0;
        for ( uint32_t lb = 0; lb < ni; lb += WaveSize )
        {
            const uint32_t l    = lb + lane;
            bool           cand = false;
            if ( l < ni )
            {
                const uint32_t j  = b.nbIds[ size_t( i ) * K + l ];
                const float4   aj = b.atoms[ j ];
                const float    Rj = aj.w + rp;
                sAtom[ w ][ l ]   = make_float4( aj.x - ci.x, aj.y - ci.y, aj.z - ci.z, Rj * Rj );
                sAbs[ w ][ l ]    = aj;
                cand              = j > i && b.status[ size_t( i ) * K + l ] == 1;
            }
            const uint64_t mask = ballot( cand );
            if ( cand )
                sCand[ w ][ m + rank_below( mask ) ] = uint16_t( l );
            m += uint32_t( __popcll( mask ) );
        }
        const bool haveTouch = K <= 128;
        if ( haveTouch && uint32_t( lane ) < min( m, TouchStride ) )
            sTouch[ w ][ lane ] = b.touch[ size_t( i ) * TouchStride + lane ];
        wave_sync();

        // ---- candidate pairs (a < c over the intersected circles of i), flattened over the lanes ----
        // Pairs are taken in rounds of PairRound: first the pairs with k in N(j) are compacted into
        // LDS in pair order (operations.inl:330-376: k must be a neighbour of j), then the geometry
        // runs on them densely, 64 at a time. Records keep pair order either way.
        const uint32_t P     = m * ( m - 1 ) / 2;
        uint32_t       nRec  = 0, nW = 0, nFp64 = 0;
        for ( uint32_t r0 = 0; r0 < P; r0 += PairRound )
        {
        const uint32_t r1 = min( P, r0 + PairRound );
        uint32_t       nMember = 0;
        for ( uint32_t pb = r0; pb < r1; pb += WaveSize )
        {
            const uint32_t p      = pb + lane;
            bool           member = false;
            uint32_t       a = 0, cc = 0;
            if ( p < r1 )
            {
                // row-major triangular index -> (a, c) with a < c
                const float fm = float( 2 * m - 1 );
                int         ai = int( ( fm - fsqrt( fm * fm - 8.f * float( p ) ) ) * 0.5f ); // corrected below
                ai             = max( 0, min( ai, int( m ) - 2 ) );
                while ( ai > 0 && uint32_t( ai ) * ( 2 * m - ai - 1 ) / 2 > p )
                    ai--;
                while ( uint32_t( ai + 1 ) * ( 2 * m - ai - 2 ) / 2 <= p )
                    ai++;
                a                 = uint32_t( ai );
                cc                = p - a * ( 2 * m - a - 1 ) / 2 + a + 1;
                const uint32_t la = sCand[ w ][ a ];
                const uint32_t lk = sCand[ w ][ cc ];
                // Unless some list was truncated at K, membership is exactly the predicate that built
                // j's list, evaluated with the same operands and operations; the slot of k in j's
                // list is found at emission.
                if ( truncated )
                {
                    const uint32_t j = b.nbIds[ size_t( i ) * K + la ], k = b.nbIds[ size_t( i ) * K + lk ];
                    member           = find_in_list( b.nbIds + size_t( j ) * K, b.nbCount[ j ], k ) >= 0;
                }
                else
                {
                    const float4 Aj = sAbs[ w ][ la ], Ak = sAbs[ w ][ lk ];
                    const f3     d  = xyz( Ak ) - xyz( Aj );
                    const float  Rs = ( Aj.w + rp ) + ( Ak.w + rp );
                    member          = len2( d ) - Rs * Rs <= PredicateEps;
                }
            }
            const uint64_t mk = ballot( member );
            if ( member )
                sPair[ w ][ nMember + rank_below( mk ) ] = a | ( cc << 16 );
            nMember += uint32_t( __popcll( mk ) );
        }
        wave_sync();
        for ( uint32_t mb = 0; mb < nMember; mb += WaveSize )
        {
            const uint32_t e    = mb + lane;
            int            nOut = 0;
            uint32_t       out0 = 0, out1 = 0;
            bool           fp64 = false;
            if ( e < nMember )
            {
                const uint32_t pk = sPair[ w ][ e ];
                const uint32_t a = pk & 0xffffu, cc = pk >> 16;
                const uint32_t la = sCand[ w ][ a ];
                const uint32_t lk = sCand[ w ][ cc ];
                {
                    // Geometry for the decisions only (hardware rcp / rsq, about 1 ulp, far inside
                    // the predicates' epsilon): the stored probe positions are solved again at fill.
                    const float4    aj  = sAtom[ w ][ la ];
                    const float4    ak  = sAtom[ w ][ lk ];
                    const f3        dj  = xyz( aj ), dk = xyz( ak );
                    const CircleGeo ij  = circle_geo_fast( dj, Ri2, aj.w );
                    const CircleGeo ik  = circle_geo_fast( dk, Ri2, ak.w );
                    const f3        u   = ik.normal - ij.normal * dot( ik.normal, ij.normal );
                    const float     den = dot( u, ik.normal );
                    const float     q   = dot( ik.centre - ij.centre, ik.normal );
                    const f3        xc  = u * ( q * frcp( den ) ) + ij.centre;
                    const float     t   = Ri2 - len2( xc );
                    f3              x[ 2 ];
                    bool            ok  = false;
                    if ( ill_conditioned( den, t, Ri2 ) )
                    {
                        fp64               = true;
                        const Resolved r64 = resolve_fp64( b.atoms[ i ], sAbs[ w ][ la ], sAbs[ w ][ lk ], rp );
                        x[ 0 ]             = r64.x0;
                        x[ 1 ]             = r64.x1;
                        ok                 = r64.ok;
                    }
                    else if ( t >= 0.f )
                    {
                        const f3 cr = cross( ik.normal, ij.normal );
                        const f3 nn = cr * ( frsq( len2( cr ) ) * fsqrt( t ) );
                        x[ 0 ]      = xc - nn;
                        x[ 1 ]      = xc + nn;
                        ok          = true;
                    }
                    if ( ok )
                    {
                        // visible iff outside every other SAS sphere of N(i) (operations.inl:415-430).
                        // A sphere containing a point of both circles touches both, so only the
                        // spheres in touch(ij) & touch(ik) can occlude (exact pruning).
                        // per probe, the least of |c_l - x|^2 - R_l^2 over the occluder candidates l
                        // (visible iff positive); integer / float state keeps the loops light
                        float mn0 = 1.f, mn1 = 1.f;
                        auto  occlude = [ & ]( uint32_t l ) {
                            const float4 al = sAtom[ w ][ l ];
                            const f3     v0 = xyz( al ) - x[ 0 ], v1 = xyz( al ) - x[ 1 ];
                            mn0 = fminf( mn0, fmaf( v0.x, v0.x, fmaf( v0.y, v0.y, fmaf( v0.z, v0.z, -al.w ) ) ) );
                            mn1 = fminf( mn1, fmaf( v1.x, v1.x, fmaf( v1.y, v1.y, fmaf( v1.z, v1.z, -al.w ) ) ) );
                        };
                        if ( haveTouch && uint32_t( a ) < TouchStride && cc < TouchStride )
                        {
                            // a touch set leaves out its circle's own sphere, so j and k are not in
                            // touch(ij) & touch(ik)
                            const uint4 ta = sTouch[ w ][ a ], tb = sTouch[ w ][ cc ];
                            uint64_t    o  = ( uint64_t( ta.y & tb.y ) << 32 ) | ( ta.x & tb.x );
                            while ( o != 0 && fmaxf( mn0, mn1 ) > 0.f )
                            {
                                occlude( uint32_t( __builtin_ctzll( o ) ) );
                                o &= o - 1;
                            }
                            o = ( uint64_t( ta.w & tb.w ) << 32 ) | ( ta.z & tb.z );
                            while ( o != 0 && fmaxf( mn0, mn1 ) > 0.f )
                            {
                                occlude( 64u + uint32_t( __builtin_ctzll( o ) ) );
                                o &= o - 1;
                            }
                        }
                        else
                            for ( uint32_t l = 0; l < ni && fmaxf( mn0, mn1 ) > 0.f; l++ )
                                if ( l != la && l != lk )
                                    occlude( l );
                        const bool vis0 = mn0 > 0.f, vis1 = mn1 > 0.f;
                        // hasNeighbors: the probe is closer than rp to the plane of the centres (:447-463)
                        const f3       pc     = cross( ij.normal, dj - dk );
                        const f3       pn     = pc * frsq( len2( pc ) );
                        const uint32_t packed = la | ( lk << 10 );
                        for ( int s = 0; s < 2; s++ )
                        {
                            if ( !( s == 0 ? vis0 : vis1 ) )
                                continue;
                            const float    d     = dot( -x[ s ], pn );
                            const bool     hasNb = d * d - rp * rp < -PredicateEps;
                            const uint32_t rec   = packed | ( s ? RecMinus : 0u ) | ( hasNb ? RecHasNb : 0u );
                            if ( nOut == 0 )
                                out0 = rec;
                            else
                                out1 = rec;
                            nOut++;
                        }
                    }
                }
            }
            const uint64_t has1 = ballot( nOut >= 1 ), has2 = ballot( nOut == 2 );
            const uint32_t off  = nRec + rank_below( has1 ) + rank_below( has2 );
            if ( nOut >= 1 && off < b.stageStride )
                stage[ off ] = out0;
            if ( nOut == 2 && off + 1 < b.stageStride )
                stage[ off + 1 ] = out1;
            nRec += uint32_t( __popcll( has1 ) + __popcll( has2 ) );
            nW += uint32_t( __popcll( ballot( nOut >= 1 && ( out0 & RecHasNb ) ) ) + __popcll( ballot( nOut == 2 && ( out1 & RecHasNb ) ) ) );
            nFp64 += uint32_t( __popcll( ballot( fp64 ) ) );
        }
        wave_sync(); // the next round overwrites the pair list
        }
        if ( nRec > b.stageStride )
        {
            // per-atom staging too small: counted, the host grows it and rebuilds
            if ( lane == 0 )
                set_overflow( b.ctl, OverflowAtomRecords );
            nRec = nW = 0;
        }
        if ( lane == 0 )
        {
            b.atomCnt[ i ]                   = nW;
            b.atomCnt[ b.n + i ]             = nRec - nW;
            b.atomCnt[ 2 * size_t( b.n ) + i ] = nFp64;
        }
    }

    // Place the staged records of atom i: with-neighbour ones at wStart[i] + rank (from the front),
    // the others at nStart[i] + rank (from the back) -- the |w|-descending order of the reference's
    // thrust::sort, made deterministic -- and count intersections per circle (complete surface).
    // A wave takes 64 consecutive atoms (about one record each) and spreads their records over its
    // lanes, 64 at a time; a record's rank within its atom is the count of same-kind records before
    // it: those of its atom earlier in the chunk (ballot) plus those of earlier chunks (LDS carry).
    constexpr int EmitWaves = 4;
    template<int MODE>
    __global__ void __launch_bounds__( EmitWaves * WaveSize ) intersections_emit_kernel( Buffers b )
    {
        __shared__ uint32_t sPre[ EmitWaves ][ WaveSize + 1 ];
        __shared__ uint32_t sCarry[ EmitWaves ][ 2 ][ WaveSize ];
        const int      w    = wave_in_block();
        const int      lane = lane_id();
        const uint32_t a0   = ( xcd_block() * EmitWaves + uint32_t( w ) ) * WaveSize;
        if ( a0 >= b.n )
            return;
        const uint32_t i0 = a0 + uint32_t( lane );
        uint32_t       wS = 0, nS = 0, cnt = 0;
        if ( i0 < b.n )
        {
            wS  = b.wStart[ i0 ];
            nS  = b.nStart[ i0 ];
            cnt = ( b.wStart[ i0 + 1 ] - wS ) + ( b.nStart[ i0 + 1 ] - nS );
        }
        const uint32_t incl  = wave_incl_scan( cnt );
        const uint32_t total = wave_bcast( incl, WaveSize - 1 );
        if ( total == 0 )
            return;
        sPre[ w ][ lane + 1 ]  = incl;
        sCarry[ w ][ 0 ][ lane ] = 0u;
        sCarry[ w ][ 1 ][ lane ] = 0u;
        if ( lane == 0 )
            sPre[ w ][ 0 ] = 0u;
        wave_sync();
        const uint32_t K  = b.K;
        const uint32_t ci = min( b.ctl->intersectedCircleNb, b.capTrimmed );
        for ( uint32_t base = 0; base < total; base += WaveSize )
        {
            const uint32_t e    = base + uint32_t( lane );
            const bool     live = e < total;
            int            t    = 0; // owning atom of record e: largest t with sPre[t] <= e
            if ( live )
            {
                int lo = 0, hi = WaveSize;
                while ( hi - lo > 1 )
                {
                    const int mid = ( lo + hi ) >> 1;
                    if ( sPre[ w ][ mid ] <= e )
                        lo = mid;
                    else
                        hi = mid;
                }
                t = lo;
            }
            const uint32_t i    = a0 + uint32_t( t );
            const uint32_t r    = e - sPre[ w ][ t ];
            uint32_t       rec  = live ? b.stage[ size_t( i ) * b.stageStride + r ] : 0u;
            const bool     isW  = live && ( rec & RecHasNb );
            const uint64_t mW = ballot( isW ), mN = ballot( live && !isW );
            // lanes of this chunk that hold earlier records of the same atom
            const uint32_t first = sPre[ w ][ t ] > base ? sPre[ w ][ t ] - base : 0u;
            const uint64_t same  = lanes_below() & ~( ( uint64_t( 1 ) << first ) - 1 );
            const uint32_t kind  = isW ? 0u : 1u;
            const uint32_t rank  = sCarry[ w ][ kind ][ t ] + uint32_t( __popcll( ( isW ? mW : mN ) & same ) );
            wave_sync();
            if ( live )
            {
                atomicAdd( &sCarry[ w ][ kind ][ t ], 1u );
                const uint32_t la = rec & 1023u, lk = ( rec >> 10 ) & 1023u;
                const uint32_t j  = b.nbIds[ size_t( i ) * K + la ];
                const uint32_t k  = b.nbIds[ size_t( i ) * K + lk ];
                // slot of k in j's ascending neighbour list (k is a member: tested at counting)
                const int      found = find_in_list( b.nbIds + size_t( j ) * K, b.nbCount[ j ], k );
                const uint32_t jk    = found >= 0 ? uint32_t( found ) : 0u;
                if ( found < 0 )
                    atomicAdd( &b.ctl->jkAnomalyNb, 1u );
                rec |= jk << 20;
                const uint32_t idx = ( isW ? b.wStart[ i ] : b.nStart[ i ] ) + rank;
                if ( idx < b.capRecords )
                    b.records[ isW ? idx : b.capRecords - 1 - idx ] = make_uint4( i, j, k, rec );
                else
                    set_overflow( b.ctl, OverflowIntersections );
                if ( MODE == 0 )
                {
                    // per-circle intersection counts on (ij, jk, ik) (operations.inl:476-482). A jk
                    // circle that is not intersected (its status rounded the other way on the j
                    // side; near-degenerate only) has no trimmed slot: the reference's popcount
                    // then names the next intersected circle of j, which depends on j's slot order,
                    // so such a patch is left out of every jk list here.
                    const bool     jkInt = is_intersected( b, j, jk );
                    const uint32_t t0 = trimmed_id( b, i, la ), t1 = jkInt ? trimmed_id( b, j, jk ) : ci;
                    const uint32_t t2 = trimmed_id( b, i, lk );
                    if ( !jkInt )
                        atomicAdd( &b.ctl->jkAnomalyNb, 1u );
                    if ( t0 < ci )
                        atomicAdd( b.circleIntNb + t0, 1u );
                    if ( t1 < ci )
                        atomicAdd( b.circleIntNb + t1, 1u );
                    if ( t2 < ci )
                        atomicAdd( b.circleIntNb + t2, 1u );
                }
            }
            wave_sync(); // carries complete before the next chunk reads them
        }
    }
    // 8-bit compare-and-swap through the enclosing aligned 32-bit word.
    __device__ __forceinline__ void promote_status( uint8_t * status, size_t slot )
    {
        uint32_t *     word  = reinterpret_cast<uint32_t *>( status + ( slot & ~size_t( 3 ) ) );
        const uint32_t shift = uint32_t( slot & 3 ) * 8;
        uint32_t       old   = __hip_atomic_load( word, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT );
        for ( ;; )
        {
            if ( ( ( old >> shift ) & 0xffu ) != 1u )
                return;
            const uint32_t nv   = ( old & ~( 0xffu << shift ) ) | ( 2u << shift );
            const uint32_t prev = atomicCAS( word, old, nv );
            if ( prev == old )
                return;
            old = prev;
        }
    }

    template<int MODE>
    __global__ void __launch_bounds__( 256 ) fill_kernel( Buffers b )
    {
        const uint32_t Iw = b.ctl->intersectionWithNbNb, In = b.ctl->intersectionNoNbNb;
        if ( Iw + In > b.capRecords || ( b.ctl->overflow & ( OverflowTrimmed | OverflowIntersections ) ) )
            return; // overflow: host grows and rebuilds
        const uint32_t I  = Iw + In;
        const float    rp = b.probeRadius;
        const uint32_t K  = b.K;
        for ( uint32_t q = blockIdx.x * blockDim.x + threadIdx.x; q < I; q += gridDim.x * blockDim.x )
        {
            const uint4    rec  = q < Iw ? b.records[ q ] : b.records[ b.capRecords - 1 - ( q - Iw ) ];
            const uint32_t i = rec.x, j = rec.y, k = rec.z;
            const int      side = ( rec.w & RecMinus ) ? -1 : 1;
            const int      hasNb = ( rec.w & RecHasNb ) ? 1 : 0;
            const uint32_t la = rec.w & 1023u, lk = ( rec.w >> 10 ) & 1023u, jk = ( rec.w >> 20 ) & 1023u;
            const float4   ai = b.atoms[ i ], aj = b.atoms[ j ], ak = b.atoms[ k ];
            const f3       ci = xyz( ai );
            const f3       dj = xyz( aj ) - ci, dk = xyz( ak ) - ci;
            const float    Ri = ai.w + rp, Rj = aj.w + rp, Rk = ak.w + rp;
            bool           ill;
            f3             x = probe_fp32( dj, dk, Ri * Ri, Rj * Rj, Rk * Rk, side, ill );
            float4 pos = make_float4( ci.x + x.x, ci.y + x.y, ci.z + x.z, float( side * ( 1 + hasNb ) ) );
            if ( ill )
            {
                // absolute position rounded once
                const d3 xd = mkd( ci ) + probe_fp64( rel64( aj, ai ), rel64( ak, ai ), ai.w, aj.w, ak.w, rp, side );
                pos.x = float( xd.x ), pos.y = float( xd.y ), pos.z = float( xd.z );
            }
            b.concavePos[ q ] = pos;
            b.concaveIds[ q ] = make_int4( int( i ), int( j ), int( k ), 0 );
            if ( MODE == 0 )
            {
                const uint32_t Ci     = min( b.ctl->intersectedCircleNb, b.capTrimmed );
                // a jk circle that is not intersected takes no patch (as at counting)
                const uint32_t ids[ 3 ] = { trimmed_id( b, i, la ), is_intersected( b, j, jk ) ? trimmed_id( b, j, jk ) : Ci,
                                            trimmed_id( b, i, lk ) };
                for ( int e = 0; e < 3; e++ )
                {
                    if ( ids[ e ] >= Ci )
                        continue;
                    const uint32_t slot = b.circleStart[ ids[ e ] ] + atomicAdd( b.circleCursor + ids[ e ], 1u );
                    if ( slot < b.capLists )
                        b.circleLists[ slot ] = q;
                    else
                        set_overflow( b.ctl, OverflowCircleLists );
                }
            }
            else
            {
                promote_status( b.status, size_t( i ) * K + la );
                promote_status( b.status, size_t( i ) * K + lk );
                promote_status( b.status, size_t( j ) * K + jk );
            }
        }
    }

    // angleBetweenEdges (sesdf/include/bcs/cuda/circles.cuh:27-31)
    __device__ __forceinline__ float angle_between_edges( f3 u, f3 v, f3 n )
    {
        const float theta = acosf( fminf( fmaxf( dot( u, v ) / ( sqrtf( len2( u ) ) * sqrtf( len2( v ) ) ), -1.f ), 1.f ) );
        const float det   = dot( u, cross( v, n ) );
        return copysignf( 1.f, det ) > 0.f ? theta : TwoPi - theta;
    }

    constexpr int SegThreads = 128;
    constexpr int SegMax     = 16;   // the reference's MaxIntersectionsPerCircles (sesdf.cu:17)
    constexpr int SegRange   = 1024; // intersected circles scanned per block and round

    // Intersected circles whose lists hold two or more patches (about one in fourteen on a globule)
    // are first gathered into an LDS queue per range of SegRange circles, then taken one per thread,
    // so that the lanes that sort lists and measure angles are dense instead of one in fourteen.
    // The output position of each circle's segments comes from segStart, not from the order.
    __global__ void __launch_bounds__( SegThreads ) segments_kernel( Buffers b )
    {
        __shared__ uint32_t sList[ SegMax ][ SegThreads ];
        __shared__ float    sAng[ SegMax ][ SegThreads ];
        __shared__ uint8_t  sOrd[ SegMax ][ SegThreads ];
        __shared__ uint32_t sQueue[ SegRange ];
        __shared__ uint32_t sQueued;
        const int      tx = threadIdx.x;
        const uint32_t Ci = min( b.ctl->intersectedCircleNb, b.capTrimmed );
        const float    rp = b.probeRadius;
        // any overflow leaves lists incomplete: skip, the host grows the buffers and rebuilds
        if ( b.ctl->overflow || b.ctl->intersectionWithNbNb + b.ctl->intersectionNoNbNb > b.capRecords ||
             b.ctl->circleListNb > b.capLists )
            return;
        for ( uint32_t r0 = blockIdx.x * uint32_t( SegRange ); r0 < Ci; r0 += gridDim.x * uint32_t( SegRange ) )
        {
            if ( tx == 0 )
                sQueued = 0;
            __syncthreads();
            const uint32_t r1 = min( r0 + uint32_t( SegRange ), Ci );
            for ( uint32_t k = r0 + uint32_t( tx ); k - uint32_t( tx ) < r1; k += SegThreads )
            {
                const bool     take = k < r1 && b.circleIntNb[ k ] / 2 != 0;
                const uint64_t m    = ballot( take );
                uint32_t       base = 0;
                if ( lane_id() == 0 && m )
                    base = atomicAdd( &sQueued, uint32_t( __popcll( m ) ) );
                base = uint32_t( __shfl( int( base ), 0, WaveSize ) );
                if ( take )
                    sQueue[ base + rank_below( m ) ] = k;
            }
            __syncthreads();
            const uint32_t queued = sQueued;
            for ( uint32_t q0 = 0; q0 < queued; q0 += SegThreads )
            {
                if ( q0 + uint32_t( tx ) >= queued )
                    continue;
                const uint32_t c   = sQueue[ q0 + tx ];
                uint32_t       cnt = b.circleIntNb[ c ];
                if ( cnt > SegMax - 1 )
                {
                    atomicAdd( &b.ctl->circleIntOverflowNb, 1u );
                    cnt = SegMax - 1;
                }
                const uint32_t g  = b.trimmedToGlobal[ c ];
                const uint32_t i  = g / b.K;
                const uint32_t j  = b.nbIds[ g ];
                const uint32_t st = b.circleStart[ c ];
                // deterministic start: the lowest record index
                for ( uint32_t q = 0; q < cnt; q++ )
                {
                    const uint32_t v = b.circleLists[ st + q ];
                    int            p = int( q ) - 1;
                    while ( p >= 0 && sList[ p ][ tx ] > v )
                    {
                        sList[ p + 1 ][ tx ] = sList[ p ][ tx ];
                        p--;
                    }
                    sList[ p + 1 ][ tx ] = v;
                }
                const float4    ai   = b.atoms[ i ], aj = b.atoms[ j ];
                const f3        ci   = xyz( ai );
                const float     Ri   = ai.w + rp, Rj = aj.w + rp;
                const CircleGeo geo  = circle_geo( xyz( aj ) - ci, Ri * Ri, Rj * Rj );
                const uint32_t  s0   = sList[ 0 ][ tx ];
                const int4      sIds = b.concaveIds[ s0 ];
                const float4    sPos = b.concavePos[ s0 ];
                const bool      secondary = i == uint32_t( sIds.x ) && j == uint32_t( sIds.z );
                const float     dir       = copysignf( 1.f, sPos.w ) * ( secondary ? -1.f : 1.f );
                const f3        ref       = ( xyz( sPos ) - ci ) - geo.centre;
                sAng[ 0 ][ tx ]           = 0.f;
                sOrd[ 0 ][ tx ]           = 0;
                for ( uint32_t q = 1; q < cnt; q++ )
                {
                    const float4 p  = b.concavePos[ sList[ q ][ tx ] ];
                    f3           v1 = ref, v2 = unit( ( xyz( p ) - ci ) - geo.centre );
                    if ( dir < 0.f )
                    {
                        const f3 tmp = v1;
                        v1           = v2;
                        v2           = tmp;
                    }
                    const float a = angle_between_edges( v1, v2, geo.normal );
                    // stable insertion by angle
                    int pidx = int( q ) - 1;
                    while ( pidx >= 0 && sAng[ pidx ][ tx ] > a )
                    {
                        sAng[ pidx + 1 ][ tx ] = sAng[ pidx ][ tx ];
                        sOrd[ pidx + 1 ][ tx ] = sOrd[ pidx ][ tx ];
                        pidx--;
                    }
                    sAng[ pidx + 1 ][ tx ] = a;
                    sOrd[ pidx + 1 ][ tx ] = uint8_t( q );
                }
                const uint32_t gs = b.segStart[ c ];
                for ( uint32_t q = 0; q < cnt / 2; q++ )
                {
                    uint32_t A = sList[ sOrd[ 2 * q ][ tx ] ][ tx ], B = sList[ sOrd[ 2 * q + 1 ][ tx ] ][ tx ];
                    if ( dir < 0.f )
                    {
                        const uint32_t tmp = A;
                        A                  = B;
                        B                  = tmp;
                    }
                    if ( gs + q < b.capSegments )
                        b.segments[ gs + q ] = make_uint4( i, j, A, B );
                    else
                        set_overflow( b.ctl, OverflowSegments );
                }
            }
            __syncthreads();
        }
    }

    // One wave per with-neighbour probe q: the rows of the cell stencil around q are flattened over
    // the lanes (the with-neighbour probes of a row are the W records of its atoms, contiguous in
    // records), every candidate is tested by one lane, and hits are appended with a ballot in
    // (row, record) order, the first maxNb kept.
    template<int MODE>
    __global__ void __launch_bounds__( 256 ) probe_neighbors_kernel( Buffers b )
    {
        __shared__ uint32_t sRowStart[ 4 ][ WaveSize ];
        __shared__ uint32_t sRowPrefix[ 4 ][ WaveSize ]; // first record of each row
        __shared__ uint32_t sRowOf[ 4 ][ WaveSize ];     // row marks of one 64-record window
        const uint32_t      Iw = b.ctl->intersectionWithNbNb;
        if ( Iw + b.ctl->intersectionNoNbNb > b.capRecords || Iw > b.capProbeNb || ( b.ctl->overflow & OverflowIntersections ) )
            return;
        const int      w     = wave_in_block();
        const int      lane  = lane_id();
        const float    rp    = b.probeRadius;
        const float    rmax  = __uint_as_float( b.ctl->radiusMaxBits );
        const float    range = 2.f * rp;
        const uint32_t maxNb = b.maxProbeNb;
        for ( uint32_t q = blockIdx.x * 4 + w; q < Iw; q += gridDim.x * 4 )
        {
        const uint32_t owner = b.records[ q ].x;
        const float4   pq    = b.concavePos[ q ];
        const Molecule M     = b.mols[ b.molNb > 1 ? b.atomMol[ owner ] : 0u ];
        // a probe within 2 rp of pq lies on the SAS sphere of its owner atom, so that atom is within
        // 2 rp + (r_max + rp) of pq
        const float search = range + rmax + rp + 1e-3f;
        const int   reach  = max( int( ceilf( search / M.cell ) ), 1 );
        int         gx, gy, gz;
        cell_of( M, pq.x, pq.y, pq.z, gx, gy, gz );
        const int   x0 = max( gx - reach, 0 ), x1 = min( gx + reach, M.dx - 1 );
        const int   y0 = max( gy - reach, 0 ), y1 = min( gy + reach, M.dy - 1 );
        const int   z0 = max( gz - reach, 0 ), z1 = min( gz + reach, M.dz - 1 );
        const int   ny = y1 - y0 + 1, rows = ny * ( z1 - z0 + 1 );
        const float wv = MODE == 0 ? pq.x * pq.y * pq.z : 0.f;
        uint32_t    count = 0;
        for ( int rowBase = 0; rowBase < rows; rowBase += WaveSize )
        {
            const int r  = rowBase + lane;
            uint32_t  st = 0, len = 0;
            if ( r < rows )
            {
                const uint32_t h0 = cell_index( M, x0, y0 + r % ny, z0 + r / ny );
                const uint32_t a0 = b.cellStart[ h0 ], a1 = b.cellStart[ h0 + uint32_t( x1 - x0 ) + 1 ];
                st                = b.wStart[ a0 ];
                len               = b.wStart[ a1 ] - st;
            }
            const uint32_t incl  = wave_incl_scan( len );
            const uint32_t excl  = incl - len;
            const uint32_t total = wave_bcast( incl, WaveSize - 1 );
            sRowStart[ w ][ lane ]  = st;
            sRowPrefix[ w ][ lane ] = excl;
            for ( uint32_t base = 0; base < total; base += WaveSize )
            {
                // row of record e = base + lane: each non-empty row marks its first record of the
                // window (the row holding record `base` marks slot 0); a prefix maximum carries
                // the marks forward
                sRowOf[ w ][ lane ] = 0u;
                wave_sync();
                if ( len > 0u && excl < base + WaveSize && incl > base )
                    sRowOf[ w ][ max( excl, base ) - base ] = uint32_t( lane );
                wave_sync();
                const int      row = wave_incl_max( int( sRowOf[ w ][ lane ] ) );
                const uint32_t e   = base + lane;
                bool           hit = false;
                float4         pr  = make_float4( 0.f, 0.f, 0.f, 0.f );
                if ( e < total )
                {
                    const uint32_t rec = sRowStart[ w ][ row ] + ( e - sRowPrefix[ w ][ row ] );
                    if ( rec != q )
                    {
                        pr  = b.concavePos[ rec ];
                        hit = sqrtf( len2( xyz( pq ) - xyz( pr ) ) ) < range;
                    }
                }
                const uint64_t mask = ballot( hit );
                const uint32_t slot = count + rank_below( mask );
                if ( hit && slot < maxNb )
                    b.concaveNb[ size_t( q ) * maxNb + slot ] = make_float4( pr.x, pr.y, pr.z, wv );
                count += uint32_t( __popcll( mask ) );
            }
            wave_sync();
        }
        if ( lane == 0 )
        {
            if ( count >= maxNb )
                atomicAdd( &b.ctl->probeNbOverflowNb, 1u );
            b.concaveIds[ q ].w = int( min( count, maxNb ) );
            if ( MODE == 1 )
                b.concavePos[ q ].w = float( q * maxNb );
        }
        }
    }

    // Exterior path: status-2 circles per atom (full, or promoted by a visible intersection), 16
    // slots per load: statuses are 0, 1, 2, so bit 1 of a byte marks status 2. One thread per atom:
    // an atom has about one torus record, so a wave per atom would leave its lanes idle.
    __device__ __forceinline__ uint32_t torus_word( const uint4 v, uint32_t q )
    {
        return q == 0 ? v.x : q == 1 ? v.y : q == 2 ? v.z : v.w;
    }
    // status-2 bits of the 4 slots [4 wd, 4 wd + 4) below ni
    __device__ __forceinline__ uint32_t torus_bits( uint32_t word, uint32_t wd, uint32_t ni )
    {
        const uint32_t valid = ni > 4 * wd ? min( ni - 4 * wd, 4u ) : 0u;
        const uint32_t keep  = valid >= 4 ? ~0u : ( 1u << ( 8 * valid ) ) - 1u;
        return word & keep & 0x02020202u;
    }

    __global__ void __launch_bounds__( 256 ) torus_count_kernel( Buffers b )
    {
        const uint32_t i = xcd_block() * blockDim.x + threadIdx.x;
        if ( i >= b.n )
            return;
        const uint32_t ni  = b.nbCount[ i ];
        const uint4 *  row = reinterpret_cast<const uint4 *>( b.status + size_t( i ) * b.K ); // K % 16 == 0
        uint32_t       cnt = 0;
        for ( uint32_t v = 0; 16 * v < ni; v++ )
        {
            const uint4 x = row[ v ];
#pragma unroll
            for ( uint32_t q = 0; q < 4; q++ )
                cnt += uint32_t( __popc( torus_bits( torus_word( x, q ), 4 * v + q, ni ) ) );
        }
        b.atomCnt[ i ] = cnt;
    }

    // Torus records (MegaMol layout) of atom i's status-2 circles, in slot order. A wave takes 64
    // atoms and spreads their records over its lanes (an atom has about one record, some several):
    // record r of the wave belongs to the atom whose exclusive prefix is the largest one <= r, and
    // is that atom's k-th status-2 slot, found by a popcount walk over its status row.
    constexpr int TorusWaves = 4;
    __global__ void __launch_bounds__( TorusWaves * WaveSize ) torus_emit_kernel( Buffers b )
    {
        __shared__ uint32_t sPre[ TorusWaves ][ WaveSize ];
        const int      w    = wave_in_block();
        const uint32_t lane = uint32_t( lane_id() );
        const uint32_t base = ( xcd_block() * TorusWaves + uint32_t( w ) ) * WaveSize;
        if ( base >= b.n )
            return;
        const uint32_t K     = b.K;
        const float    rp    = b.probeRadius;
        const uint32_t last  = min( base + WaveSize, b.n );
        const uint32_t i0    = base + lane;
        const uint32_t cnt   = i0 < last ? b.torusBase[ i0 + 1 ] - b.torusBase[ i0 ] : 0u;
        const uint32_t incl  = wave_incl_scan( cnt );
        const uint32_t total = wave_bcast( incl, WaveSize - 1 );
        const uint32_t first = b.torusBase[ base ];
        sPre[ w ][ lane ]    = incl - cnt;
        wave_sync();
        for ( uint32_t r0 = 0; r0 < total; r0 += WaveSize )
        {
            const uint32_t r = r0 + lane;
            if ( r >= total )
                break;
            // atom slot: the last one whose exclusive prefix is <= r (empty atoms share prefixes)
            int lo = 0, hi = WaveSize;
            while ( hi - lo > 1 )
            {
                const int mid = ( lo + hi ) >> 1;
                if ( sPre[ w ][ mid ] <= r )
                    lo = mid;
                else
                    hi = mid;
            }
            const uint32_t i   = base + uint32_t( lo );
            uint32_t       k   = r - sPre[ w ][ lo ];
            const uint32_t dst = first + r;
            if ( dst >= b.capTorus )
            {
                set_overflow( b.ctl, OverflowTorus );
                continue;
            }
            const uint32_t ni  = b.nbCount[ i ];
            const uint4 *  row = reinterpret_cast<const uint4 *>( b.status + size_t( i ) * K );
            uint32_t       l   = 0;
            for ( uint32_t v = 0; 16 * v < ni; v++ )
            {
                const uint4 x     = row[ v ];
                bool        found = false;
                for ( uint32_t q = 0; q < 4 && !found; q++ )
                {
                    uint32_t       bits = torus_bits( torus_word( x, q ), 4 * v + q, ni );
                    const uint32_t c    = uint32_t( __popc( bits ) );
                    if ( k < c )
                    {
                        for ( uint32_t e = 0; e < k; e++ )
                            bits &= bits - 1;
                        l     = 4 * ( 4 * v + q ) + uint32_t( __ffs( int( bits ) ) - 1 ) / 8;
                        found = true;
                    }
                    else
                        k -= c;
                }
                if ( found )
                    break;
            }
            const float4    ai    = b.atoms[ i ];
            const f3        ci    = xyz( ai );
            const float     Ri    = ai.w + rp;
            const uint32_t  j     = b.nbIds[ size_t( i ) * K + l ];
            const float4    aj    = b.atoms[ j ];
            const f3        dj    = xyz( aj ) - ci;
            const float     Rj    = aj.w + rp;
            const CircleGeo geo   = circle_geo( dj, Ri * Ri, Rj * Rj );
            const float     rad   = sqrtf( geo.radius2 );
            const f3        ta    = unit( geo.centre );
            const f3        ortho = unit( cross( ta, mk( 0.f, 0.f, 1.f ) ) );
            const f3        P     = geo.centre + ortho * rad;
            const f3        X     = unit( P ) * ai.w;
            const float     lpi   = sqrtf( len2( P ) );
            const f3        C     = dj * ( lpi / ( sqrtf( len2( P - dj ) ) + lpi ) );
            const float     dist  = sqrtf( len2( X - C ) );
            const f3        Cout  = C - geo.centre;
            b.torusPos[ dst ]     = make_float4( ci.x + geo.centre.x, ci.y + geo.centre.y, ci.z + geo.centre.z, rad );
            b.torusAxis[ dst ]    = make_float4( ta.x, ta.y, ta.z, rp );
            b.torusVs[ dst ]      = make_float4( Cout.x, Cout.y, Cout.z, dist );
        }
    }

    static uint32_t persistent_blocks( uint64_t work, uint32_t threads, uint32_t perCu )
    {
        const uint64_t wanted = ( work + threads - 1 ) / threads;
        const uint64_t cap    = uint64_t( device_cu_count() ) * perCu;
        const uint64_t g      = wanted < cap ? wanted : cap;
        return uint32_t( g > 0 ? g : 1 );
    }

    void launch_intersections( const Buffers & b, hipStream_t s )
    {
        if ( b.n == 0 )
            return;
        // block shape (waves per workgroup) chosen by the autotuner
        const uint32_t wv = b.wavesInter == 2 || b.wavesInter == 8 ? b.wavesInter : 4u;
        const uint32_t cb = ( b.n + wv - 1 ) / wv;
        if ( b.K <= 128 )
        {
            if ( wv == 2 )
                intersections_count_kernel<128, 2><<<cb, 2 * WaveSize, 0, s>>>( b );
            else if ( wv == 8 )
                intersections_count_kernel<128, 8><<<cb, 8 * WaveSize, 0, s>>>( b );
            else
                intersections_count_kernel<128, 4><<<cb, 4 * WaveSize, 0, s>>>( b );
        }
        else
            intersections_count_kernel<320, 4><<<( b.n + 3 ) / 4, 4 * WaveSize, 0, s>>>( b );
    }

    void launch_intersections_finish( const Buffers & b, hipStream_t s )
    {
        if ( b.n == 0 )
            return;
        ScanArgs       a {};
        a.n           = b.n;
        a.ticket      = TicketIntersections;
        a.stage       = 3;
        a.in[ 0 ]     = b.atomCnt;
        a.in[ 1 ]     = b.atomCnt + b.n;
        a.in[ 2 ]     = b.atomCnt + 2 * size_t( b.n ); // FP64 re-solve counts (summed, not scanned)
        a.out[ 0 ]    = b.wStart;
        a.out[ 1 ]    = b.nStart;
        a.totals[ 0 ] = &b.ctl->intersectionWithNbNb;
        a.totals[ 1 ] = &b.ctl->intersectionNoNbNb;
        a.maxOut      = &b.ctl->maxAtomRecords;
        a.sumOut      = &b.ctl->fp64ResolveNb;
        launch_scan<2>( b, a, s );
        const uint32_t emitBlocks = ( b.n + EmitWaves * WaveSize - 1 ) / ( EmitWaves * WaveSize );
        if ( b.mode == 0 )
            intersections_emit_kernel<0><<<emitBlocks, EmitWaves * WaveSize, 0, s>>>( b );
        else
            intersections_emit_kernel<1><<<emitBlocks, EmitWaves * WaveSize, 0, s>>>( b );
    }

    // Per-circle list offsets (exclusive scan of the intersection counts) and segment offsets
    // (scan of count / 2) over the Ci intersected circles; clears the list cursors.
    void launch_circle_scan( const Buffers & b, hipStream_t s )
    {
        if ( b.n == 0 || b.mode != 0 )
            return;
        ScanArgs a {};
        a.n           = b.capTrimmed;
        a.countPtr    = &b.ctl->intersectedCircleNb;
        a.ticket      = TicketCircleScan;
        a.stage       = 4;
        a.in[ 0 ]     = b.circleIntNb;
        a.halfSecond  = 1;
        a.out[ 0 ]    = b.circleStart;
        a.out[ 1 ]    = b.segStart;
        a.totals[ 0 ] = &b.ctl->circleListNb;
        a.totals[ 1 ] = &b.ctl->segmentNb;
        a.clearAlso   = b.circleCursor;
        launch_scan<2>( b, a, s );
    }

    void launch_fill( const Buffers & b, hipStream_t s )
    {
        if ( b.n == 0 )
            return;
        const uint32_t blocks = persistent_blocks( b.capRecords, 256, 8 );
        if ( b.mode == 0 )
            fill_kernel<0><<<blocks, 256, 0, s>>>( b );
        else
            fill_kernel<1><<<blocks, 256, 0, s>>>( b );
    }

    void launch_segments( const Buffers & b, hipStream_t s )
    {
        if ( b.n == 0 || b.mode != 0 )
            return;
        const uint32_t blocks = persistent_blocks( b.capTrimmed, SegRange, 8 );
        segments_kernel<<<blocks, SegThreads, 0, s>>>( b );
    }

    void launch_probe_neighbors( const Buffers & b, hipStream_t s )
    {
        if ( b.n == 0 )
            return;
        // one wave per with-neighbour probe, striding beyond 64 workgroups per CU
        const uint32_t blocks = persistent_blocks( uint64_t( b.capProbeNb ) * WaveSize, 256, 64 );
        if ( b.mode == 0 )
            probe_neighbors_kernel<0><<<blocks, 256, 0, s>>>( b );
        else
            probe_neighbors_kernel<1><<<blocks, 256, 0, s>>>( b );
    }

    void launch_torus( const Buffers & b, hipStream_t s )
    {
        if ( b.n == 0 || b.mode != 1 )
            return;
        const uint32_t blocks = ( b.n + 255 ) / 256;
        torus_count_kernel<<<blocks, 256, 0, s>>>( b );
        ScanArgs a {};
        a.n           = b.n;
        a.ticket      = TicketTorus;
        a.stage       = 5;
        a.in[ 0 ]     = b.atomCnt;
        a.out[ 0 ]    = b.torusBase;
        a.totals[ 0 ] = &b.ctl->torusNb;
        launch_scan<1>( b, a, s );
        torus_emit_kernel<<<( b.n + TorusWaves * WaveSize - 1 ) / ( TorusWaves * WaveSize ), TorusWaves * WaveSize, 0, s>>>( b );
    }
} // namespace sesdf
