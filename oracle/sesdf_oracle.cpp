// ORACLE — CPU restatement of PlathC/SESDF's solvent-excluded-surface construction.
//
// TEST INFRASTRUCTURE ONLY. This file is the checker for the HIP product path: only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it. The product library never
// links it and never falls back to it.
//
// It restates, stage by stage and in the reference's order and predicates, the host classes
// bcs::Sesdf::build() (sesdf/src/sesdf/sesdf.cu:172-264, complete surface) and
// bcs::Ssesdf::build() (sesdf/src/ssesdf/ssesdf.cu:133-187, exterior path), templated on the
// scalar type:
//   * T = float  : plain IEEE FP32 (no fast-math emulation) -- the host C++ baseline;
//   * T = double : the accuracy oracle. With tracking on, every threshold predicate is also
//                  classified robust / ambiguous against an epsilon band (|value - threshold| <= delta)
//                  so that GPU-vs-oracle topology differences can be attributed to near-degenerate
//                  cases and counted.
// All indices are in the reference's grid-sorted atom order; sortedToInitial maps them back.
//
// Parity pins (tests/test_oracle.py, tests/test_oracle_sdf.py): brute-force O(N^2)/O(N^3) enumeration on small inputs,
// closed forms for 1/2/3/4-atom configurations, and the invariants sum(per-circle counts) = 3I,
// even per-circle counts, G = 3I/2, status(ij) = status(ji), i<j<k in every concave id.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <string>
#include <thread>
#include <type_traits>
#include <vector>

namespace orc
{
    template<class T>
    struct V3
    {
        T x, y, z;
    };
    template<class T> inline V3<T> operator+( V3<T> a, V3<T> b ) { return { a.x + b.x, a.y + b.y, a.z + b.z }; }
    template<class T> inline V3<T> operator-( V3<T> a, V3<T> b ) { return { a.x - b.x, a.y - b.y, a.z - b.z }; }
    template<class T> inline V3<T> operator*( V3<T> a, T s ) { return { a.x * s, a.y * s, a.z * s }; }
    template<class T> inline V3<T> operator*( T s, V3<T> a ) { return { s * a.x, s * a.y, s * a.z }; }
    template<class T> inline V3<T> operator/( V3<T> a, T s ) { return { a.x / s, a.y / s, a.z / s }; }
    template<class T> inline T     dot( V3<T> a, V3<T> b ) { return a.x * b.x + a.y * b.y + a.z * b.z; }
    template<class T> inline T     len2( V3<T> a ) { return dot( a, a ); }
    template<class T> inline T     len( V3<T> a ) { return std::sqrt( dot( a, a ) ); }
    template<class T> inline V3<T> cross( V3<T> a, V3<T> b )
    {
        return { a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x };
    }
    // helper_math normalize: v * rsqrt(dot(v, v))
    template<class T> inline V3<T> normalize( V3<T> v ) { return v * ( T( 1 ) / std::sqrt( dot( v, v ) ) ); }
    template<class T> inline T     sgn( T f ) { return std::copysign( T( 1 ), f ); } // math.cuh:13-16
    // helper_math.h clamp: fmaxf(lo, fminf(v, hi)) -- a NaN argument yields hi
    template<class T> inline T     clampv( T v, T lo, T hi ) { return std::fmax( lo, std::fmin( v, hi ) ); }
    template<class T> inline T     det3( V3<T> a, V3<T> b, V3<T> c ) { return dot( a, cross( b, c ) ); } // math.cuh:27-30

    struct AtomF
    {
        float x, y, z, r;
    };

    // bcs::GridInfo (sesdf/include/bcs/cuda/grid.cuh:8-17, grid.inl:8-28)
    struct GridInfo
    {
        float ox, oy, oz;
        float cx, cy, cz;
        int   size;

        void pos( float px, float py, float pz, int & gx, int & gy, int & gz ) const
        {
            gx = static_cast<int>( std::floor( ( px - ox ) / cx ) );
            gy = static_cast<int>( std::floor( ( py - oy ) / cy ) );
            gz = static_cast<int>( std::floor( ( pz - oz ) / cz ) );
        }
        uint32_t hash( int gx, int gy, int gz ) const
        {
            gx &= ( size - 1 );
            gy &= ( size - 1 );
            gz &= ( size - 1 );
            return uint32_t( gz * size * size + gy * size + gx );
        }
    };

    constexpr float    MaxVdwRadius = 3.48f;            // sesdf.cu:23, operations.cu:80
    constexpr float    Eps          = 1e-4f;            // predicate slack used throughout the reference
    constexpr uint32_t Empty        = 0xffffffffu;      // AccelerationGrid::EmptyGridCellValue
    constexpr float    TwoPif       = 6.283185307f;     // core/math.hpp

    // nextPowerOfTwoValue(nextPowerOfTwoExponent(n)) (core/math.inl:5-19)
    inline uint32_t referenceGridSize( uint32_t n )
    {
        uint32_t e = 0, b = n;
        while ( b >>= 1 )
            e++;
        uint32_t v = 1;
        while ( e > v )
            v <<= 1;
        return v;
    }

    // Constructor grid configuration (sesdf.cu:23-33) from the vdW AABB (molecule.cpp:132-143).
    inline GridInfo referenceGrid( const AtomF * atoms, uint32_t n, float rp )
    {
        float mn[ 3 ] = { 3.4e38f, 3.4e38f, 3.4e38f }, mx[ 3 ] = { -3.4e38f, -3.4e38f, -3.4e38f };
        for ( uint32_t i = 0; i < n; i++ )
        {
            const float p[ 3 ] = { atoms[ i ].x, atoms[ i ].y, atoms[ i ].z };
            for ( int d = 0; d < 3; d++ )
            {
                mn[ d ] = std::min( p[ d ] - atoms[ i ].r, mn[ d ] );
                mx[ d ] = std::max( p[ d ] + atoms[ i ].r, mx[ d ] );
            }
        }
        GridInfo       g;
        const uint32_t gs = referenceGridSize( n );
        float          o[ 3 ], c[ 3 ];
        for ( int d = 0; d < 3; d++ )
        {
            o[ d ]           = mn[ d ] - MaxVdwRadius - rp;
            const float size = std::fabs( mx[ d ] + MaxVdwRadius + rp - o[ d ] );
            c[ d ]           = size / float( gs );
        }
        g.ox = o[ 0 ], g.oy = o[ 1 ], g.oz = o[ 2 ];
        g.cx = c[ 0 ], g.cy = c[ 1 ], g.cz = c[ 2 ];
        g.size = int( gs );
        return g;
    }

    // AccelerationGrid::build (grid.cu:99-132): hash, stable sort by hash, cell start/end.
    struct SortedGrid
    {
        std::vector<uint32_t> sortedIdx;  // sorted -> initial
        std::vector<uint32_t> cellStart, cellEnd;
    };

    template<class P>
    SortedGrid buildGrid( const GridInfo & g, uint32_t n, P posOf )
    {
        SortedGrid            sg;
        std::vector<uint32_t> hashes( n );
        for ( uint32_t i = 0; i < n; i++ )
        {
            float x, y, z;
            posOf( i, x, y, z );
            int gx, gy, gz;
            g.pos( x, y, z, gx, gy, gz );
            hashes[ i ] = g.hash( gx, gy, gz );
        }
        sg.sortedIdx.resize( n );
        for ( uint32_t i = 0; i < n; i++ )
            sg.sortedIdx[ i ] = i;
        std::stable_sort( sg.sortedIdx.begin(), sg.sortedIdx.end(), [ & ]( uint32_t a, uint32_t b ) { return hashes[ a ] < hashes[ b ]; } );
        const size_t cells = size_t( g.size ) * g.size * g.size;
        sg.cellStart.assign( cells, Empty );
        sg.cellEnd.assign( cells, 0 );
        for ( uint32_t s = 0; s < n; s++ )
        {
            const uint32_t h = hashes[ sg.sortedIdx[ s ] ];
            if ( s == 0 || h != hashes[ sg.sortedIdx[ s - 1 ] ] )
                sg.cellStart[ h ] = s;
            sg.cellEnd[ h ] = s + 1;
        }
        return sg;
    }

    template<class Fn>
    void parallelFor( uint32_t n, int threads, Fn fn )
    {
        if ( threads <= 1 || n < 256 )
        {
            for ( uint32_t i = 0; i < n; i++ )
                fn( i );
            return;
        }
        std::atomic<uint32_t>    next { 0 };
        std::vector<std::thread> pool;
        for ( int t = 0; t < threads; t++ )
            pool.emplace_back(
                [ & ]
                {
                    for ( ;; )
                    {
                        const uint32_t b = next.fetch_add( 64 );
                        if ( b >= n )
                            break;
                        const uint32_t e = std::min( n, b + 64 );
                        for ( uint32_t i = b; i < e; i++ )
                            fn( i );
                    }
                } );
        for ( auto & th : pool )
            th.join();
    }

    // Output of one build, mirroring SesdfData / SsesdfData (sesdf/include/bcs/sesdf/graphics.hpp:39-66,
    // ssesdf/graphics.hpp:31-50) plus the internal arrays the reference exposes through getters.
    struct Result
    {
        int      mode = 0; // 0 complete (Sesdf), 1 exterior path (Ssesdf)
        uint32_t n = 0, K = 0;
        float    rp = 0;
        GridInfo grid {};

        std::vector<float>    atoms;            // sorted float4
        std::vector<uint32_t> sortedToInitial;  // n
        std::vector<uint32_t> nbCount;          // n
        std::vector<uint32_t> nbIds;            // n*K
        std::vector<uint8_t>  status;           // n*K (final statuses, after exterior promotion)
        std::vector<uint32_t> trimmedToGlobal;  // Ci (complete)
        std::vector<uint32_t> convexPatches;    // uint2 n (complete)
        std::vector<float>    sectors;          // float4 V (complete)
        std::vector<uint32_t> fullCircles;      // uint2 Cf (complete)
        std::vector<int32_t>  concaveIds;       // int4 I
        std::vector<float>    concavePos;       // float4 I (FP32-rounded)
        std::vector<double>   concavePos64;     // xyz I at T precision
        std::vector<float>    concaveNb;        // float4 32*Iw
        std::vector<uint32_t> circleIntNb;      // per trimmed circle intersection count (complete)
        std::vector<uint32_t> segments;         // uint4 G (complete)
        std::vector<float>    torus;            // 3 x float4 T (exterior): pos[T], axis[T], vs[T]
        std::vector<uint32_t> torusSlots;       // global slot id per torus record
        uint32_t              Iw = 0;

        // epsilon-band tracking (double precision with track = true)
        std::vector<uint32_t> ambPairs;       // (i, j) sorted idx, i < j: neighbour predicate within delta
        std::vector<uint8_t>  ambCircle;      // n*K: status decision within delta
        std::vector<uint8_t>  ambConcave;     // I: existence/side ambiguous
        std::vector<uint8_t>  ambHasNb;       // I: hasNeighbors predicate within delta
        std::vector<uint32_t> ambRejected;    // (i, j, k, side) candidates rejected within delta
        std::vector<uint8_t>  ambSegCircle;   // per trimmed circle: its segments are ambiguous
        std::vector<uint8_t>  ambProbeNb;     // Iw: probe neighbour set ambiguous
        uint64_t              anomalyJk = 0;  // records whose jk circle is not status 1
        uint64_t              overflowNb = 0, overflowProbeNb = 0, overflowCircleInt = 0;
        std::vector<uint64_t> stats;          // predicate evaluation counts (visibility tests, occluder tests, ...)
    };

    template<class T>
    struct Geometry
    {
        V3<T> center, normal;
        T     radius;
    };

    // SesdfContext::getCircleGeometry (sesdf/include/bcs/sesdf/data.cuh:98-114) with
    // getCircleCenter (cuda/circles.cuh:9-14).
    template<class T>
    inline Geometry<T> circleGeometry( V3<T> ci, T Ri, V3<T> cj, T Rj )
    {
        const T     d = len2( ci - cj );
        const T     t = ( Ri * Ri - Rj * Rj + d ) / ( T( 2 ) * d );
        Geometry<T> g;
        g.center                = ci + ( cj - ci ) * t;
        const V3<T> circleToI   = ci - g.center;
        const T     sqDist      = dot( circleToI, circleToI );
        g.radius                = std::sqrt( Ri * Ri - sqDist );
        g.normal                = normalize( cj - ci );
        return g;
    }

    // bcs::orthogonalVector (cuda/math.cuh:32-38)
    template<class T>
    inline V3<T> orthogonalVector( V3<T> n )
    {
        V3<T> ref { 1, 0, 0 };
        if ( std::fabs( dot( n, ref ) ) > T( 1 ) - T( 1e-4f ) )
            ref = { 0, 1, 0 };
        return normalize( cross( n, ref ) );
    }

    // bcs::angleBetweenEdges (cuda/circles.cuh:27-31)
    template<class T>
    inline T angleBetweenEdges( V3<T> u, V3<T> v, V3<T> n )
    {
        const T theta = std::acos( clampv( dot( u, v ) / ( len( u ) * len( v ) ), T( -1 ), T( 1 ) ) );
        return sgn( det3( u, v, n ) ) > T( 0 ) ? theta : T( TwoPif ) - theta;
    }

    struct Record
    {
        uint32_t i, a, j, b, k, jkIdx;
        int      side;      // +1: x1 = xc - n, -1: x2 = xc + n
        int      hasNb;
        uint8_t  amb, ambHas;
    };

    template<class T>
    struct Builder
    {
        const AtomF *      input;
        uint32_t           n;
        T                  rp;
        float              rpf;
        uint32_t           K;
        int                mode;
        bool               track;
        double             delta, deltaTheta;
        int                threads;
        Result &           res;
        // Optional neighbour lists to use in place of the grid query's (per input atom: a count and
        // K input ids). Test infrastructure: lets a build whose lists were truncated at K in another
        // grid order be checked stage by stage on exactly those lists.
        const uint32_t *   nbOverrideCount = nullptr;
        const uint32_t *   nbOverrideIds   = nullptr;
        // Optional sorted order (sorted index -> input index) to use in place of the grid's: the atom
        // that heads a triple (lowest sorted index) is the one whose list bounds its probe tests, so
        // builds with truncated lists are comparable only in the same order.
        const uint32_t *   orderOverride   = nullptr;

        void fillSorted()
        {
            res.atoms.resize( size_t( n ) * 4 );
            c.resize( n ), r.resize( n ), R.resize( n );
            for ( uint32_t s = 0; s < n; s++ )
            {
                const AtomF & a = input[ res.sortedToInitial[ s ] ];
                res.atoms[ 4 * s + 0 ] = a.x, res.atoms[ 4 * s + 1 ] = a.y, res.atoms[ 4 * s + 2 ] = a.z, res.atoms[ 4 * s + 3 ] = a.r;
                c[ s ] = { T( a.x ), T( a.y ), T( a.z ) };
                r[ s ] = T( a.r );
                R[ s ] = T( a.r ) + rp;
            }
        }

        // Renames every sorted index produced by the grid query into the order `orderOverride`.
        void reorder()
        {
            std::vector<uint32_t> newOfInput( n ), oldToNew( n );
            for ( uint32_t s = 0; s < n; s++ )
                newOfInput[ orderOverride[ s ] ] = s;
            for ( uint32_t s = 0; s < n; s++ )
                oldToNew[ s ] = newOfInput[ res.sortedToInitial[ s ] ];
            std::vector<uint32_t> cnt( n ), ids( size_t( n ) * K, 0 );
            for ( uint32_t s = 0; s < n; s++ )
            {
                const uint32_t t = oldToNew[ s ];
                cnt[ t ]         = res.nbCount[ s ];
                for ( uint32_t l = 0; l < cnt[ t ]; l++ )
                    ids[ size_t( t ) * K + l ] = oldToNew[ res.nbIds[ size_t( s ) * K + l ] ];
                std::sort( ids.begin() + size_t( t ) * K, ids.begin() + size_t( t ) * K + cnt[ t ] );
            }
            res.nbCount.swap( cnt );
            res.nbIds.swap( ids );
            for ( size_t q = 0; q + 1 < res.ambPairs.size(); q += 2 )
            {
                const uint32_t a = oldToNew[ res.ambPairs[ q ] ], b = oldToNew[ res.ambPairs[ q + 1 ] ];
                res.ambPairs[ q ] = std::min( a, b ), res.ambPairs[ q + 1 ] = std::max( a, b );
            }
            res.sortedToInitial.assign( orderOverride, orderOverride + n );
            fillSorted();
        }

        std::vector<V3<T>> c;   // sorted centres
        std::vector<T>     r;   // sorted vdW radii
        std::vector<T>     R;   // r + rp

        Builder( const AtomF * in, uint32_t n_, float rp_, uint32_t K_, int mode_, bool track_, double d_, double dt_, int th, Result & out ) :
            input( in ), n( n_ ), rp( T( rp_ ) ), rpf( rp_ ), K( K_ ), mode( mode_ ), track( track_ ), delta( d_ ), deltaTheta( dt_ ),
            threads( th ), res( out )
        {
        }

        bool inBand( T v ) const { return track && std::fabs( double( v ) ) <= delta; }

        // -- #2 findCircles: grid, sort, neighbours (operations.inl:14-30, operations.cu:34-124) ----------
        void neighbours()
        {
            res.grid       = referenceGrid( input, n, rpf );
            SortedGrid sg  = buildGrid( res.grid, n, [ & ]( uint32_t i, float & x, float & y, float & z ) { x = input[ i ].x, y = input[ i ].y, z = input[ i ].z; } );
            res.sortedToInitial = sg.sortedIdx;
            cellStart_          = sg.cellStart;
            cellEnd_            = sg.cellEnd;
            fillSorted();
            res.nbCount.assign( n, 0 );
            res.nbIds.assign( size_t( n ) * K, 0 );
            std::vector<std::vector<uint32_t>> amb( threads > 1 ? n : 0 );
            std::vector<uint32_t>              ambSeq;
            std::atomic<uint64_t>              overflow { 0 };
            const GridInfo &                   g = res.grid;
            parallelFor( n, threads, [ & ]( uint32_t i ) {
                const AtomF a { res.atoms[ 4 * i ], res.atoms[ 4 * i + 1 ], res.atoms[ 4 * i + 2 ], res.atoms[ 4 * i + 3 ] };
                int         gx, gy, gz;
                g.pos( a.x, a.y, a.z, gx, gy, gz );
                const float range = a.r + MaxVdwRadius + 2.f * rpf;
                const int   rx = std::max( int( std::ceil( range / g.cx ) ), 1 );
                const int   ry = std::max( int( std::ceil( range / g.cy ) ), 1 );
                const int   rz = std::max( int( std::ceil( range / g.cz ) ), 1 );
                const int   sx = std::max( gx - rx, 0 ), sy = std::max( gy - ry, 0 ), sz = std::max( gz - rz, 0 );
                const int   ex = std::min( gx + rx, g.size - 1 ), ey = std::min( gy + ry, g.size - 1 ), ez = std::min( gz + rz, g.size - 1 );
                uint32_t    count = 0;
                bool        full  = false;
                for ( int z = sz; z <= ez && !full; z++ )
                    for ( int y = sy; y <= ey && !full; y++ )
                        for ( int x = sx; x <= ex && !full; x++ )
                        {
                            const uint32_t h  = g.hash( x, y, z );
                            const uint32_t st = sgCellStart( h );
                            if ( st == Empty )
                                continue;
                            const uint32_t en = sgCellEnd( h );
                            for ( uint32_t j = st; j < en; j++ )
                            {
                                if ( j == i )
                                    continue;
                                const T d = len2( c[ i ] - c[ j ] );
                                const T e = R[ i ] + R[ j ];
                                const T m = d - e * e;
                                if ( track && j > i && inBand( m - T( Eps ) ) )
                                {
                                    if ( threads > 1 )
                                        amb[ i ].push_back( j );
                                    else
                                        ambSeq.push_back( i ), ambSeq.push_back( j );
                                }
                                if ( m <= T( Eps ) )
                                {
                                    res.nbIds[ size_t( i ) * K + count ] = j;
                                    count++;
                                    if ( count == K )
                                    {
                                        overflow++;
                                        full = true;
                                        break;
                                    }
                                }
                            }
                        }
                res.nbCount[ i ] = count;
            } );
            res.overflowNb = overflow;
            if ( orderOverride )
                reorder(); // the atom grid is not used past this point
            if ( nbOverrideCount )
            {
                std::vector<uint32_t> initialToSorted( n );
                for ( uint32_t s = 0; s < n; s++ )
                    initialToSorted[ res.sortedToInitial[ s ] ] = s;
                for ( uint32_t i = 0; i < n; i++ )
                {
                    const uint32_t a   = res.sortedToInitial[ i ];
                    const uint32_t cnt = std::min( nbOverrideCount[ a ], K );
                    uint32_t *     out = res.nbIds.data() + size_t( i ) * K;
                    for ( uint32_t l = 0; l < cnt; l++ )
                        out[ l ] = initialToSorted[ nbOverrideIds[ size_t( a ) * K + l ] ];
                    std::sort( out, out + cnt ); // ascending, as the grid traversal writes them
                    res.nbCount[ i ] = cnt;
                }
            }
            if ( threads > 1 )
            {
                for ( uint32_t i = 0; i < n; i++ )
                    for ( uint32_t j : amb[ i ] )
                        res.ambPairs.push_back( i ), res.ambPairs.push_back( j );
            }
            else
                res.ambPairs = ambSeq;
        }
        // grid lookups bound after neighbours() builds the grid
        std::vector<uint32_t> cellStart_, cellEnd_;
        uint32_t              sgCellStart( uint32_t h ) const { return cellStart_[ h ]; }
        uint32_t              sgCellEnd( uint32_t h ) const { return cellEnd_[ h ]; }

        // -- computeCirclesVisibilityStatus (sesdf operations.cu:126-204; ssesdf operations.cu:169-244) ------
        void visibility()
        {
            res.status.assign( size_t( n ) * K, 0 );
            res.ambCircle.assign( track ? size_t( n ) * K : 0, 0 );
            std::atomic<uint64_t> tests { 0 };
            parallelFor( n, threads, [ & ]( uint32_t i ) {
                const uint32_t ni    = res.nbCount[ i ];
                uint64_t       local = 0;
                for ( uint32_t a = 0; a < ni; a++ )
                {
                    const uint32_t j = res.nbIds[ size_t( i ) * K + a ];
                    if ( mode == 1 && !( i < j ) )
                        continue; // exterior path: only i < j is evaluated (ssesdf operations.cu:200-204)
                    const Geometry<T> g       = circleGeometry( c[ i ], R[ i ], c[ j ], R[ j ] );
                    uint32_t          visible = 1, complete = 1;
                    bool              hiddenRobust = false, hiddenAmb = false, nearRobust = false, nearAmb = false;
                    // No circle where the SAS spheres do not meet in one: one contains the other
                    // (|c_i - c_j| < |R_i - R_j|) or they are tangent or apart (|c_i - c_j| >=
                    // R_i + R_j, kept by the neighbour test's slack). The reference's radius is then
                    // the square root of a negative number, NaN, which passes no test (status 2
                    // with NaN geometry); defined here as hidden. In band within delta of either.
                    {
                        const T d2 = len2( c[ j ] - c[ i ] ), dR = R[ i ] - R[ j ], dS = R[ i ] + R[ j ];
                        const bool edge = inBand( d2 - dR * dR ) || inBand( d2 - dS * dS );
                        if ( !( d2 > dR * dR && d2 < dS * dS ) )
                        {
                            res.status[ size_t( i ) * K + a ] = 0;
                            if ( track )
                                res.ambCircle[ size_t( i ) * K + a ] = edge;
                            continue;
                        }
                        if ( track && edge )
                            hiddenAmb = true;
                    }
                    for ( uint32_t kk = 0; kk < ni; kk++ )
                    {
                        if ( kk == a )
                            continue;
                        const uint32_t k  = res.nbIds[ size_t( i ) * K + kk ];
                        const T        Rk = R[ k ];
                        local++;
                        const V3<T> kToCircle      = g.center - c[ k ];
                        const T     distToCircle   = len( kToCircle );
                        const T     secondDistance = dot( kToCircle, g.normal );
                        // The reference divides by distToCircle and takes sqrt(1 - cos^2) as is: a
                        // sphere centred on the circle's centre (0/0) or on its axis (|cos| rounding
                        // above 1) gives NaN there, and NaN compares as "neither hides nor touches".
                        // Here both take their continuous limit (far = near = radius^2 at the centre),
                        // as the HIP kernels' h^2 / A form does.
                        const T     cosTheta       = distToCircle > T( 0 ) ? secondDistance / distToCircle : T( 1 );
                        const T     sinTheta       = std::sqrt( std::max( T( 0 ), T( 1 ) - cosTheta * cosTheta ) );
                        const T     firstFar       = sinTheta * distToCircle + g.radius;
                        const T     far            = firstFar * firstFar + secondDistance * secondDistance;
                        const T     mFar           = far - Rk * Rk - T( Eps );
                        const T     firstNear      = ( -sinTheta ) * distToCircle + g.radius;
                        const T     nearv          = firstNear * firstNear + secondDistance * secondDistance;
                        const T     mNear          = nearv - Rk * Rk - T( Eps );
                        if ( track )
                        {
                            // evaluate every k so that the decision can be classified robust / ambiguous
                            if ( inBand( mFar ) )
                                hiddenAmb = true;
                            else if ( mFar < 0 )
                                hiddenRobust = true;
                            if ( inBand( mNear ) )
                                nearAmb = true;
                            else if ( mNear < 0 )
                                nearRobust = true;
                        }
                        if ( mFar < 0 )
                        {
                            visible = 0;
                            if ( !track )
                                break;
                        }
                        if ( complete && mNear < 0 )
                            complete = 0;
                    }
                    res.status[ size_t( i ) * K + a ] = uint8_t( visible * ( 1 + complete ) );
                    if ( track )
                    {
                        // hidden decided robustly if some k hides robustly; otherwise ambiguous if any k is in band.
                        bool amb = false;
                        if ( !hiddenRobust )
                        {
                            if ( hiddenAmb )
                                amb = true;
                            else if ( !nearRobust && nearAmb )
                                amb = true;
                        }
                        res.ambCircle[ size_t( i ) * K + a ] = amb;
                    }
                }
                tests += local;
            } );
            res.stats.push_back( tests );
        }

        // -- trimmed ids, sectors, full circles (operations.inl:32-93, 175-228) ---------------------------
        std::vector<uint32_t> globalToTrimmedScan;  // exclusive scan of the intersected mask, n*K+1
        void compaction()
        {
            if ( mode != 0 )
                return;
            const size_t total = size_t( n ) * K;
            globalToTrimmedScan.assign( total + 1, 0 );
            uint32_t acc = 0;
            for ( size_t g = 0; g < total; g++ )
            {
                globalToTrimmedScan[ g ] = acc;
                const uint32_t i = uint32_t( g / K ), a = uint32_t( g % K );
                if ( a < res.nbCount[ i ] && res.status[ g ] == 1 && res.nbIds[ g ] > i )
                {
                    res.trimmedToGlobal.push_back( uint32_t( g ) );
                    acc++;
                }
            }
            globalToTrimmedScan[ total ] = acc;
            res.convexPatches.resize( size_t( n ) * 2 );
            uint32_t written = 0;
            for ( uint32_t i = 0; i < n; i++ )
            {
                res.convexPatches[ 2 * i ] = written;
                for ( uint32_t a = 0; a < res.nbCount[ i ]; a++ )
                {
                    const size_t   g = size_t( i ) * K + a;
                    const uint32_t j = res.nbIds[ g ];
                    if ( res.status[ g ] == 0 )
                        continue;
                    const T           rirp   = R[ i ];
                    const V3<T>       cij    = circleGeometry( c[ i ], R[ i ], c[ j ], R[ j ] ).center;
                    const V3<T>       cijci  = c[ i ] - cij;
                    const T           radius = std::sqrt( rirp * rirp - dot( cijci, cijci ) );
                    const V3<T>       cicj   = normalize( c[ j ] - c[ i ] );
                    const V3<T>       nn     = orthogonalVector( cicj );
                    const V3<T>       x      = cij + nn * radius;
                    const T           ang    = std::acos( clampv( dot( cicj, ( x - c[ i ] ) / rirp ), T( -1 ), T( 1 ) ) );
                    res.sectors.push_back( float( cicj.x ) );
                    res.sectors.push_back( float( cicj.y ) );
                    res.sectors.push_back( float( cicj.z ) );
                    res.sectors.push_back( float( ang ) );
                    written++;
                    if ( res.status[ g ] == 2 && j > i )
                        res.fullCircles.push_back( i ), res.fullCircles.push_back( j );
                }
                res.convexPatches[ 2 * i + 1 ] = written;
            }
        }

        int findInNeighbors( uint32_t j, uint32_t k ) const
        {
            const uint32_t nj = res.nbCount[ j ];
            for ( uint32_t q = 0; q < nj; q++ )
                if ( res.nbIds[ size_t( j ) * K + q ] == k )
                    return int( q );
            return -1;
        }
        bool pairAmbiguous( uint32_t a, uint32_t b ) const
        {
            if ( !track )
                return false;
            const uint32_t lo = std::min( a, b ), hi = std::max( a, b );
            return std::binary_search( ambPairSorted.begin(), ambPairSorted.end(), ( uint64_t( lo ) << 32 ) | hi );
        }
        std::vector<uint64_t> ambPairSorted;

        // -- findIntersectionsBetweenCircles (sesdf operations.inl:230-486; ssesdf operation.inl:100-351) ------
        std::vector<Record> records;
        void intersections()
        {
            if ( track )
            {
                for ( size_t q = 0; q + 1 < res.ambPairs.size(); q += 2 )
                    ambPairSorted.push_back( ( uint64_t( res.ambPairs[ q ] ) << 32 ) | res.ambPairs[ q + 1 ] );
                std::sort( ambPairSorted.begin(), ambPairSorted.end() );
            }
            std::vector<std::vector<Record>>   perAtom( n );
            std::vector<std::vector<uint32_t>> rejected( track ? n : 0 );
            std::atomic<uint64_t>              candTests { 0 }, occTests { 0 };
            parallelFor( n, threads, [ & ]( uint32_t i ) {
                const uint32_t ni = res.nbCount[ i ];
                auto           isCand = [ & ]( uint32_t a, bool & amb ) {
                    const size_t   g = size_t( i ) * K + a;
                    const uint32_t j = res.nbIds[ g ];
                    amb              = track && res.ambCircle[ g ];
                    return j > i && ( res.status[ g ] == 1 || amb );
                };
                uint64_t ct = 0, ot = 0;
                for ( uint32_t a = 0; a < ni; a++ )
                {
                    bool ambA;
                    if ( !isCand( a, ambA ) )
                        continue;
                    const uint32_t j     = res.nbIds[ size_t( i ) * K + a ];
                    const bool     realA = res.status[ size_t( i ) * K + a ] == 1;
                    for ( uint32_t b = 0; b < ni; b++ )
                    {
                        bool ambB;
                        if ( !isCand( b, ambB ) )
                            continue;
                        const uint32_t k = res.nbIds[ size_t( i ) * K + b ];
                        if ( !( j < k ) )
                            continue;
                        const bool realB  = res.status[ size_t( i ) * K + b ] == 1;
                        const int  jkIdx  = findInNeighbors( j, k );
                        const bool ambJk  = pairAmbiguous( j, k );
                        const bool real   = realA && realB && jkIdx >= 0;
                        const bool ambAny = ambA || ambB || ambJk || pairAmbiguous( i, j ) || pairAmbiguous( i, k );
                        if ( !real && !ambAny )
                            continue;
                        ct++;
                        const Geometry<T> ij = circleGeometry( c[ i ], R[ i ], c[ j ], R[ j ] );
                        const Geometry<T> ik = circleGeometry( c[ i ], R[ i ], c[ k ], R[ k ] );
                        const V3<T>       u  = ik.normal - dot( ik.normal, ij.normal ) * ij.normal;
                        const V3<T>       xc = ij.center + ( u * dot( ik.center - ij.center, ik.normal ) ) / dot( u, ik.normal );
                        const T           t  = -len2( xc - c[ i ] ) + R[ i ] * R[ i ];
                        const bool        tAmb = inBand( t );
                        if ( !( t >= T( 0 ) ) )
                        {
                            if ( tAmb )
                                for ( int side : { 1, -1 } )
                                    rejected[ i ].insert( rejected[ i ].end(), { i, j, k, uint32_t( side ) } );
                            continue;
                        }
                        const V3<T> nn = std::sqrt( t ) * normalize( cross( ik.normal, ij.normal ) );
                        const V3<T> xs[ 2 ] = { xc - nn, xc + nn };
                        bool        vis[ 2 ] = { true, true }, amb[ 2 ] = { false, false }, robustHidden[ 2 ] = { false, false };
                        for ( uint32_t l = 0; l < ni; l++ )
                        {
                            if ( l == a || l == b )
                                continue;
                            const uint32_t li = res.nbIds[ size_t( i ) * K + l ];
                            const T        Rl2 = R[ li ] * R[ li ];
                            for ( int s = 0; s < 2; s++ )
                            {
                                ot++;
                                const T m = len2( c[ li ] - xs[ s ] ) - Rl2;
                                if ( inBand( m ) )
                                    amb[ s ] = true;
                                else if ( m < 0 )
                                    robustHidden[ s ] = true;
                                vis[ s ] = vis[ s ] && m > T( 0 );
                            }
                            if ( !track && !vis[ 0 ] && !vis[ 1 ] )
                                break;
                        }
                        const V3<T> p2p3 = normalize( c[ j ] - c[ i ] );
                        const V3<T> p4p3 = normalize( c[ j ] - c[ k ] );
                        const V3<T> pn   = normalize( cross( p2p3, p4p3 ) );
                        for ( int s = 0; s < 2; s++ )
                        {
                            const bool sideAmb = ambAny || tAmb || ( amb[ s ] && !robustHidden[ s ] );
                            if ( real && vis[ s ] )
                            {
                                const T d      = dot( c[ i ] - xs[ s ], pn );
                                const T hm     = d * d - rp * rp + T( Eps );
                                Record  rec;
                                rec.i = i, rec.a = a, rec.j = j, rec.b = b, rec.k = k, rec.jkIdx = uint32_t( jkIdx );
                                rec.side   = s == 0 ? 1 : -1;
                                rec.hasNb  = hm < T( 0 );
                                rec.amb    = sideAmb;
                                rec.ambHas = inBand( hm );
                                perAtom[ i ].push_back( rec );
                            }
                            else if ( track && sideAmb && !robustHidden[ s ] )
                                rejected[ i ].insert( rejected[ i ].end(), { i, j, k, uint32_t( s == 0 ? 1 : -1 ) } );
                        }
                    }
                }
                candTests += ct;
                occTests += ot;
            } );
            res.stats.push_back( candTests );
            res.stats.push_back( occTests );
            // thrust::sort with HasNeighborsComparator (|w| descending) -- made deterministic here:
            // with-neighbour probes first, each bucket in (i, ij slot, ik slot, side) order.
            for ( int pass = 0; pass < 2; pass++ )
                for ( uint32_t i = 0; i < n; i++ )
                    for ( const Record & rec : perAtom[ i ] )
                        if ( ( pass == 0 ) == ( rec.hasNb == 1 ) )
                            records.push_back( rec );
            res.Iw = 0;
            for ( const Record & rec : records )
                res.Iw += rec.hasNb;
            if ( track )
                for ( uint32_t i = 0; i < n; i++ )
                    res.ambRejected.insert( res.ambRejected.end(), rejected[ i ].begin(), rejected[ i ].end() );
        }

        // -- fillIntersections (sesdf operations.cu:206-269; ssesdf operations.cu:287-329) --------------
        V3<T> probePosition( uint32_t i, uint32_t j, uint32_t k, int side ) const
        {
            const Geometry<T> ij = circleGeometry( c[ i ], R[ i ], c[ j ], R[ j ] );
            const Geometry<T> ik = circleGeometry( c[ i ], R[ i ], c[ k ], R[ k ] );
            const V3<T>       u  = ik.normal - dot( ik.normal, ij.normal ) * ij.normal;
            const T           t  = dot( ik.center - ij.center, ik.normal ) / dot( u, ik.normal );
            const V3<T>       x1 = ij.center + t * u;
            const T           sq = len2( x1 - c[ i ] );
            const T           cc = std::max( -sq + R[ i ] * R[ i ], T( 0 ) );
            const V3<T>       nn = normalize( cross( ij.normal, ik.normal ) );
            return x1 + ( T( side ) * std::sqrt( cc ) ) * nn;
        }

        std::vector<uint32_t> circleLists, circleStart; // complete: per trimmed circle intersection lists
        std::vector<uint32_t> recCircles;                // 3 circle ids per record (trimmed or global)
        void fill()
        {
            const uint32_t I = uint32_t( records.size() );
            res.concaveIds.resize( size_t( I ) * 4 );
            res.concavePos.resize( size_t( I ) * 4 );
            res.concavePos64.resize( size_t( I ) * 3 );
            res.ambConcave.resize( I );
            res.ambHasNb.resize( I );
            recCircles.resize( size_t( I ) * 3 );
            const size_t total = size_t( n ) * K;
            for ( uint32_t q = 0; q < I; q++ )
            {
                const Record & rec = records[ q ];
                const int      w   = rec.side * ( 1 + rec.hasNb );
                const V3<T>    x   = probePosition( rec.i, rec.j, rec.k, rec.side );
                res.concavePos[ 4 * q + 0 ] = float( x.x );
                res.concavePos[ 4 * q + 1 ] = float( x.y );
                res.concavePos[ 4 * q + 2 ] = float( x.z );
                res.concavePos[ 4 * q + 3 ] = float( w );
                res.concavePos64[ 3 * q + 0 ] = double( x.x );
                res.concavePos64[ 3 * q + 1 ] = double( x.y );
                res.concavePos64[ 3 * q + 2 ] = double( x.z );
                res.concaveIds[ 4 * q + 0 ] = int32_t( rec.i );
                res.concaveIds[ 4 * q + 1 ] = int32_t( rec.j );
                res.concaveIds[ 4 * q + 2 ] = int32_t( rec.k );
                res.concaveIds[ 4 * q + 3 ] = 0;
                res.ambConcave[ q ] = rec.amb;
                res.ambHasNb[ q ]   = rec.ambHas;
                const size_t gij = size_t( rec.i ) * K + rec.a, gik = size_t( rec.i ) * K + rec.b, gjk = size_t( rec.j ) * K + rec.jkIdx;
                if ( mode == 0 )
                {
                    // ids = (trimmed ij, trimmed jk, trimmed ik) -- operations.inl:437-440
                    recCircles[ 3 * q + 0 ] = globalToTrimmedScan[ gij ];
                    recCircles[ 3 * q + 1 ] = globalToTrimmedScan[ gjk ];
                    recCircles[ 3 * q + 2 ] = globalToTrimmedScan[ gik ];
                    if ( res.status[ gjk ] != 1 )
                    {
                        // The jk circle is not intersected (its status rounded the other way on
                        // j's side; near-degenerate only). The reference's exclusive scan then
                        // names j's next intersected circle, which depends on the order of j's
                        // neighbour slots; the patch is kept out of every jk list instead, and
                        // the record is ambiguous.
                        res.anomalyJk++;
                        recCircles[ 3 * q + 1 ] = 0xffffffffu;
                        if ( track )
                            res.ambConcave[ q ] = 1;
                    }
                }
                else
                {
                    // ids = (global ij, global ik, global jk) -- ssesdf operation.inl:308-310
                    recCircles[ 3 * q + 0 ] = uint32_t( gij );
                    recCircles[ 3 * q + 1 ] = uint32_t( gik );
                    recCircles[ 3 * q + 2 ] = uint32_t( gjk );
                    (void)total;
                }
            }
            if ( mode == 0 )
            {
                const uint32_t Ci = uint32_t( res.trimmedToGlobal.size() );
                res.circleIntNb.assign( Ci, 0 );
                for ( uint32_t q = 0; q < I; q++ )
                    for ( int e = 0; e < 3; e++ )
                        if ( recCircles[ 3 * q + e ] < Ci )
                            res.circleIntNb[ recCircles[ 3 * q + e ] ]++;
                circleStart.assign( Ci + 1, 0 );
                for ( uint32_t t = 0; t < Ci; t++ )
                    circleStart[ t + 1 ] = circleStart[ t ] + res.circleIntNb[ t ];
                circleLists.assign( circleStart[ Ci ], 0 );
                std::vector<uint32_t> fillc( Ci, 0 );
                for ( uint32_t q = 0; q < I; q++ )
                    for ( int e = 0; e < 3; e++ )
                    {
                        const uint32_t t = recCircles[ 3 * q + e ];
                        if ( t < Ci )
                            circleLists[ circleStart[ t ] + fillc[ t ]++ ] = q;
                    }
            }
            else
            {
                // atomicCAS status 1 -> 2 on the three circles (ssesdf operations.cu:301-303)
                for ( uint32_t q = 0; q < I; q++ )
                    for ( int e = 0; e < 3; e++ )
                    {
                        uint8_t & s = res.status[ recCircles[ 3 * q + e ] ];
                        if ( s == 1 )
                            s = 2;
                    }
            }
        }

        // -- buildConicPatches / buildSegments (operations.inl:488-668) ----------------------------------
        void segmentsStage()
        {
            if ( mode != 0 )
                return;
            const uint32_t Ci = uint32_t( res.trimmedToGlobal.size() );
            res.ambSegCircle.assign( Ci, 0 );
            // circles touched by an ambiguous candidate are ambiguous
            if ( track )
            {
                for ( uint32_t q = 0; q < records.size(); q++ )
                    if ( res.ambConcave[ q ] )
                        for ( int e = 0; e < 3; e++ )
                            if ( recCircles[ 3 * q + e ] < Ci )
                                res.ambSegCircle[ recCircles[ 3 * q + e ] ] = 1;
                markRejectedCircles();
            }
            for ( uint32_t t = 0; t < Ci; t++ )
            {
                uint32_t cnt = res.circleIntNb[ t ];
                if ( cnt / 2 == 0 )
                    continue;
                if ( cnt % 2 )
                    res.ambSegCircle[ t ] = 1;
                if ( cnt > 15 )
                {
                    res.overflowCircleInt++;
                    cnt = 15;
                }
                const uint32_t g  = res.trimmedToGlobal[ t ];
                const uint32_t ci = g / K, cj = res.nbIds[ g ];
                const Geometry<T> geo = circleGeometry( c[ ci ], R[ ci ], c[ cj ], R[ cj ] );
                const uint32_t *  ids = circleLists.data() + circleStart[ t ];
                const uint32_t    st  = ids[ 0 ];
                const bool        isSecondary = ci == uint32_t( res.concaveIds[ 4 * st ] ) && cj == uint32_t( res.concaveIds[ 4 * st + 2 ] );
                const T           dir = sgn( T( res.concavePos[ 4 * st + 3 ] ) ) * ( isSecondary ? T( -1 ) : T( 1 ) );
                const V3<T>       startPos { T( res.concavePos64[ 3 * st ] ), T( res.concavePos64[ 3 * st + 1 ] ), T( res.concavePos64[ 3 * st + 2 ] ) };
                const V3<T>       ref = startPos - geo.center;
                std::vector<std::pair<T, uint32_t>> ang;
                ang.push_back( { T( 0 ), 0u } );
                for ( uint32_t q = 1; q < cnt; q++ )
                {
                    const uint32_t id = ids[ q ];
                    const V3<T>    p { T( res.concavePos64[ 3 * id ] ), T( res.concavePos64[ 3 * id + 1 ] ), T( res.concavePos64[ 3 * id + 2 ] ) };
                    V3<T>          v1 = ref, v2 = normalize( p - geo.center );
                    if ( dir < T( 0 ) )
                        std::swap( v1, v2 );
                    ang.push_back( { angleBetweenEdges( v1, v2, geo.normal ), q } );
                }
                std::stable_sort( ang.begin(), ang.end(), []( auto & x, auto & y ) { return x.first < y.first; } );
                if ( track )
                {
                    for ( size_t q = 0; q < ang.size(); q++ )
                    {
                        const double a0 = double( ang[ q ].first );
                        const double a1 = q + 1 < ang.size() ? double( ang[ q + 1 ].first ) : 2 * M_PI;
                        if ( a1 - a0 <= deltaTheta )
                            res.ambSegCircle[ t ] = 1;
                    }
                }
                for ( uint32_t q = 0; q < cnt / 2; q++ )
                {
                    uint32_t A = ids[ ang[ 2 * q ].second ], B = ids[ ang[ 2 * q + 1 ].second ];
                    if ( dir < T( 0 ) )
                        std::swap( A, B );
                    res.segments.insert( res.segments.end(), { ci, cj, A, B } );
                }
            }
        }
        void markRejectedCircles()
        {
            // a rejected-but-ambiguous candidate (i, j, k) may exist on the GPU: its circles' segments are ambiguous
            std::map<uint64_t, uint32_t> slotToTrimmed;
            for ( uint32_t t = 0; t < res.trimmedToGlobal.size(); t++ )
            {
                const uint32_t g = res.trimmedToGlobal[ t ];
                const uint32_t i = g / K, j = res.nbIds[ g ];
                slotToTrimmed[ ( uint64_t( i ) << 32 ) | j ] = t;
            }
            auto mark = [ & ]( uint32_t a, uint32_t b ) {
                auto it = slotToTrimmed.find( ( uint64_t( std::min( a, b ) ) << 32 ) | std::max( a, b ) );
                if ( it != slotToTrimmed.end() )
                    res.ambSegCircle[ it->second ] = 1;
            };
            for ( size_t q = 0; q + 3 < res.ambRejected.size(); q += 4 )
            {
                const uint32_t i = res.ambRejected[ q ], j = res.ambRejected[ q + 1 ], k = res.ambRejected[ q + 2 ];
                mark( i, j ), mark( i, k ), mark( j, k );
            }
            // a circle whose status is ambiguous may gain or lose segments
            for ( uint32_t t = 0; t < res.trimmedToGlobal.size(); t++ )
                if ( res.ambCircle[ res.trimmedToGlobal[ t ] ] )
                    res.ambSegCircle[ t ] = 1;
        }

        // -- handleIntersectionSingularities + findIntersectionsNeighbors (sesdf operations.cu:14-32, 271-338;
        //    ssesdf operations.cu:57-75, 331-399) ------------------------------------------------------------
        void probeNeighbours()
        {
            const uint32_t Iw = res.Iw;
            const uint32_t M  = 32;
            res.concaveNb.assign( size_t( Iw ) * M * 4, 0.f );
            res.ambProbeNb.assign( Iw, 0 );
            if ( Iw == 0 )
                return;
            const GridInfo & g  = res.grid;
            SortedGrid       pg = buildGrid( g, Iw, [ & ]( uint32_t q, float & x, float & y, float & z ) {
                x = res.concavePos[ 4 * q ], y = res.concavePos[ 4 * q + 1 ], z = res.concavePos[ 4 * q + 2 ];
            } );
            const float range = 2.f * rpf;
            for ( uint32_t s = 0; s < Iw; s++ )
            {
                const uint32_t orig = pg.sortedIdx[ s ];
                const float    px = res.concavePos[ 4 * orig ], py = res.concavePos[ 4 * orig + 1 ], pz = res.concavePos[ 4 * orig + 2 ];
                const V3<T>    p64 { T( res.concavePos64[ 3 * orig ] ), T( res.concavePos64[ 3 * orig + 1 ] ), T( res.concavePos64[ 3 * orig + 2 ] ) };
                int            gx, gy, gz;
                g.pos( px, py, pz, gx, gy, gz );
                const int rx = std::max( int( std::ceil( range / g.cx ) ), 1 );
                const int ry = std::max( int( std::ceil( range / g.cy ) ), 1 );
                const int rz = std::max( int( std::ceil( range / g.cz ) ), 1 );
                const int sx = std::max( gx - rx, 0 ), sy = std::max( gy - ry, 0 ), sz = std::max( gz - rz, 0 );
                const int ex = std::min( gx + rx, g.size - 1 ), ey = std::min( gy + ry, g.size - 1 ), ez = std::min( gz + rz, g.size - 1 );
                uint32_t  count = 0;
                bool      full  = false;
                for ( int z = sz; z <= ez && !full; z++ )
                    for ( int y = sy; y <= ey && !full; y++ )
                        for ( int x = sx; x <= ex && !full; x++ )
                        {
                            const uint32_t h  = g.hash( x, y, z );
                            const uint32_t st = pg.cellStart[ h ];
                            if ( st == Empty )
                                continue;
                            for ( uint32_t t = st; t < pg.cellEnd[ h ]; t++ )
                            {
                                if ( t == s )
                                    continue;
                                const uint32_t o2 = pg.sortedIdx[ t ];
                                const V3<T>    q64 { T( res.concavePos64[ 3 * o2 ] ), T( res.concavePos64[ 3 * o2 + 1 ] ), T( res.concavePos64[ 3 * o2 + 2 ] ) };
                                const T        dist = len( p64 - q64 );
                                if ( inBand( dist - T( range ) ) )
                                    res.ambProbeNb[ orig ] = 1;
                                if ( dist < T( range ) )
                                {
                                    float * o = &res.concaveNb[ ( size_t( orig ) * M + count ) * 4 ];
                                    o[ 0 ] = res.concavePos[ 4 * o2 ], o[ 1 ] = res.concavePos[ 4 * o2 + 1 ], o[ 2 ] = res.concavePos[ 4 * o2 + 2 ];
                                    o[ 3 ] = mode == 0 ? px * py * pz : 0.f;
                                    if ( res.ambConcave[ o2 ] || res.ambHasNb[ o2 ] )
                                        res.ambProbeNb[ orig ] = 1;
                                    count++;
                                    if ( count == M )
                                    {
                                        res.overflowProbeNb++;
                                        full = true;
                                        break;
                                    }
                                }
                            }
                        }
                res.concaveIds[ 4 * orig + 3 ] = int32_t( count );
                if ( mode == 1 )
                    res.concavePos[ 4 * orig + 3 ] = float( orig * M );
            }
        }

        // -- fillCircles (ssesdf operations.cu:246-285; operation.inl:73-95) --------------------------------
        void torusStage()
        {
            if ( mode != 1 )
                return;
            std::vector<float> pos, axis, vs;
            for ( uint32_t i = 0; i < n; i++ )
                for ( uint32_t a = 0; a < K; a++ )
                {
                    const size_t g = size_t( i ) * K + a;
                    if ( a >= res.nbCount[ i ] || res.status[ g ] != 2 )
                        continue;
                    const uint32_t    j   = res.nbIds[ g ];
                    const Geometry<T> geo = circleGeometry( c[ i ], R[ i ], c[ j ], R[ j ] );
                    const V3<T>       ta  = normalize( geo.center - c[ i ] );
                    const V3<T>       tc  = geo.center;
                    const V3<T>       ortho = normalize( cross( ta, V3<T> { 0, 0, 1 } ) );
                    const V3<T>       P   = tc + ( ortho * geo.radius );
                    const V3<T>       X   = normalize( P - c[ i ] ) * r[ i ];
                    V3<T>             C   = ( len( P - c[ i ] ) / ( len( P - c[ j ] ) + len( P - c[ i ] ) ) ) * ( c[ j ] - c[ i ] );
                    const T           distance = len( X - C );
                    C                      = ( C + c[ i ] ) - tc;
                    pos.insert( pos.end(), { float( tc.x ), float( tc.y ), float( tc.z ), float( geo.radius ) } );
                    axis.insert( axis.end(), { float( ta.x ), float( ta.y ), float( ta.z ), rpf } );
                    vs.insert( vs.end(), { float( C.x ), float( C.y ), float( C.z ), float( distance ) } );
                    res.torusSlots.push_back( uint32_t( g ) );
                }
            res.torus = pos;
            res.torus.insert( res.torus.end(), axis.begin(), axis.end() );
            res.torus.insert( res.torus.end(), vs.begin(), vs.end() );
        }

        void run()
        {
            neighbours();
            visibility();
            compaction();
            intersections();
            fill();
            segmentsStage();
            probeNeighbours();
            torusStage();
        }
    };
    // ------------------------------------------------------------------------------------------------
    // SDF oracle (FP64). The reference ships no SDF evaluator; this library defines
    //   SDF(q) = clamp(rp + f(q), -B, B),  f = signed distance to the boundary of the SAS union A
    // (negative inside), evaluated over the trimmed boundary elements of the complete surface:
    //   faces    : sphere a, when the radial projection of q onto it is not inside another sphere;
    //   arcs     : full circles and segments (A -> B counter-clockwise about n(i -> j));
    //   vertices : concave patch (probe) positions.
    // Outside A, f = min_a |q - c_a| - R_a. The zero level set is the SES; the field is 1-Lipschitz.
    // Exterior variant: only elements of the boundary component that holds the point of largest x
    // count (components are formed over arcs, vertices and faces, see buildComponents), and points
    // outside A whose region of the complement is bounded by another component (probe-accessible
    // cavities) are inside.
    // Winner codes: type << 40 | index; types 0 outside(atom), 1 face(atom), 2 full circle,
    // 3 segment, 4 vertex, 5 clamped.
    // ------------------------------------------------------------------------------------------------
    struct SdfOracle
    {
        const Result &        r;
        double                rp;
        std::vector<V3<double>> c;
        std::vector<double>   R;
        std::vector<uint32_t> fullStart, segStart, vtxStart, vtxList;
        uint32_t              extComp = 0;
        // uniform grid over atoms
        double                cell = 1, ox = 0, oy = 0, oz = 0;
        int                   dx = 1, dy = 1, dz = 1;
        std::vector<uint32_t> cellStart, cellAtoms;
        double                rmax = 0;

        explicit SdfOracle( const Result & res ) : r( res ), rp( double( res.rp ) )
        {
            const uint32_t n = r.n;
            c.resize( n ), R.resize( n );
            for ( uint32_t a = 0; a < n; a++ )
            {
                c[ a ] = { double( r.atoms[ 4 * a ] ), double( r.atoms[ 4 * a + 1 ] ), double( r.atoms[ 4 * a + 2 ] ) };
                R[ a ] = double( r.atoms[ 4 * a + 3 ] ) + rp;
                rmax   = std::max( rmax, double( r.atoms[ 4 * a + 3 ] ) );
            }
            // element ranges per owner atom (lowest sorted index)
            auto ranges = [ & ]( std::vector<uint32_t> & start, const std::vector<uint32_t> & owner ) {
                start.assign( n + 1, 0 );
                for ( uint32_t o : owner )
                    start[ o + 1 ]++;
                for ( uint32_t a = 0; a < n; a++ )
                    start[ a + 1 ] += start[ a ];
            };
            std::vector<uint32_t> fo, so;
            for ( size_t q = 0; q + 1 < r.fullCircles.size(); q += 2 )
                fo.push_back( r.fullCircles[ q ] );
            for ( size_t q = 0; q + 3 < r.segments.size(); q += 4 )
                so.push_back( r.segments[ q ] );
            ranges( fullStart, fo );
            ranges( segStart, so );
            // vertices by owner i, in record order
            const size_t I = r.concaveIds.size() / 4;
            std::vector<uint32_t> vo( I );
            for ( size_t q = 0; q < I; q++ )
                vo[ q ] = uint32_t( r.concaveIds[ 4 * q ] );
            ranges( vtxStart, vo );
            vtxList.resize( I );
            std::vector<uint32_t> fillv( vtxStart.begin(), vtxStart.end() - 1 );
            for ( size_t q = 0; q < I; q++ )
                vtxList[ fillv[ vo[ q ] ]++ ] = uint32_t( q );
            buildComponents();
        }

        // ---- surface components (exterior variant) -------------------------------------------------
        // The boundary of A is a closed surface made of faces (exposed regions of a sphere), arcs and
        // vertices; each of its connected pieces bounds one connected region of the complement of A.
        // Components are formed over arcs (full circles, segments): a segment joins its two end
        // vertices, and the boundary loops of one face (an atom's exposed region may be split into
        // several faces, each bounded by one or more loops) are joined. Loops of atom a bound the same
        // face iff a path on the sphere between points just inside them crosses every loop of the
        // atom an even number of times. The exterior component holds the point of largest x of the
        // atom with the largest x + r (ties: lowest sorted index).
        //   atomLabel[a]: component of atom a's boundary when it is one (an atom without arcs is its
        //                 own component if exposed), MixedLabel when it spans more, NoneLabel if buried.
        static constexpr uint32_t MixedLabel = 0xfffffffeu, NoneLabel = 0xffffffffu;
        uint32_t              Cf = 0, G = 0, I = 0;
        std::vector<uint32_t> arcLabel, vtxLabel, atomLabel;
        std::vector<uint32_t> incStart, incArcs; // arcs incident to each atom (arc id: f < Cf, Cf + segment)

        uint32_t arcU( uint32_t e ) const { return e < Cf ? r.fullCircles[ 2 * e ] : r.segments[ 4 * ( e - Cf ) ]; }
        uint32_t arcV( uint32_t e ) const { return e < Cf ? r.fullCircles[ 2 * e + 1 ] : r.segments[ 4 * ( e - Cf ) + 1 ]; }
        uint32_t arcOther( uint32_t e, uint32_t a ) const { return arcU( e ) == a ? arcV( e ) : arcU( e ); }
        // cap of circle (a, o) on sphere a: pole (unit, towards c_o) and cos of its half angle
        void capOf( uint32_t a, uint32_t o, V3<double> & pole, double & cosT ) const
        {
            const V3<double> d  = c[ o ] - c[ a ];
            const double     dd = len( d );
            pole                = d * ( 1.0 / dd );
            cosT                = ( R[ a ] * R[ a ] - R[ o ] * R[ o ] + dd * dd ) / ( 2 * dd * R[ a ] );
        }
        // x (world, on circle (i, j) of the arc) within the arc: segments run from A counter-clockwise
        // about n(i -> j) to B
        bool onArc( uint32_t e, V3<double> x ) const
        {
            if ( e < Cf )
                return true;
            const uint32_t * sg = &r.segments[ 4 * size_t( e - Cf ) ];
            V3<double>       ctr, nrm;
            double           rad;
            circle( sg[ 0 ], sg[ 1 ], ctr, nrm, rad );
            const V3<double> A = pos64( sg[ 2 ] ) - ctr, B = pos64( sg[ 3 ] ) - ctr, X = x - ctr;
            auto an = [ & ]( V3<double> u ) {
                double t = std::atan2( dot( cross( A, u ), nrm ), dot( A, u ) );
                return t < 0 ? t + 2 * M_PI : t;
            };
            return an( X ) <= an( B );
        }
        // angles phi in [0, Phi] where the geodesic v(phi) = cos phi v1 + sin phi w (v1 -> v2 on
        // sphere a) meets arc e; returns the count, writes up to 2 angles
        int crossArc( uint32_t a, V3<double> v1, V3<double> w, double Phi, uint32_t e, double * phis ) const
        {
            V3<double> pole;
            double     cosT;
            capOf( a, arcOther( e, a ), pole, cosT );
            const double p = dot( v1, pole ), q = dot( w, pole ), rho = std::hypot( p, q );
            if ( !( rho > std::fabs( cosT ) ) )
                return 0;
            const double ph0 = std::atan2( q, p ), dph = std::acos( std::min( std::max( cosT / rho, -1.0 ), 1.0 ) );
            int          cnt = 0;
            for ( double ph : { ph0 - dph, ph0 + dph } )
            {
                ph = std::fmod( ph + 4 * M_PI, 2 * M_PI );
                if ( ph > Phi )
                    continue;
                const V3<double> v = v1 * std::cos( ph ) + w * std::sin( ph );
                if ( onArc( e, c[ a ] + v * R[ a ] ) )
                    phis[ cnt++ ] = ph;
            }
            return cnt;
        }
        // geodesic frame from v1 towards v2 (unit vectors); false if they are (nearly) antipodal
        static bool frame( V3<double> v1, V3<double> v2, V3<double> & w, double & Phi )
        {
            const V3<double> m = cross( v1, v2 );
            const double     s = len( m );
            if ( s < 1e-9 && dot( v1, v2 ) < 0 )
                return false;
            const V3<double> t = v2 - v1 * dot( v1, v2 );
            const double     tl = len( t );
            w                   = tl > 0 ? t * ( 1.0 / tl ) : V3<double> { 0, 0, 0 };
            Phi                 = std::atan2( s, dot( v1, v2 ) );
            return true;
        }
        // arc crossed first by the geodesic from u (an exposed point of sphere a) towards the pole
        // of one of a's caps (a buried point): it bounds the face holding u. NoneLabel if none.
        uint32_t firstHit( uint32_t a, V3<double> u ) const
        {
            for ( uint32_t t = incStart[ a ]; t < incStart[ a + 1 ]; t++ )
            {
                V3<double> pole, w;
                double     cosT, Phi;
                capOf( a, arcOther( incArcs[ t ], a ), pole, cosT );
                if ( !frame( u, pole, w, Phi ) )
                    continue;
                double   bestPhi = 1e300;
                uint32_t bestArc = NoneLabel;
                for ( uint32_t t2 = incStart[ a ]; t2 < incStart[ a + 1 ]; t2++ )
                {
                    double   ph[ 2 ];
                    const int k = crossArc( a, u, w, Phi, incArcs[ t2 ], ph );
                    for ( int q = 0; q < k; q++ )
                        if ( ph[ q ] < bestPhi )
                            bestPhi = ph[ q ], bestArc = incArcs[ t2 ];
                }
                return bestArc;
            }
            return NoneLabel;
        }
        // first arc of atom a crossed by the geodesic from u towards target (unit vectors)
        uint32_t firstHitTowards( uint32_t a, V3<double> u, V3<double> target ) const
        {
            V3<double> w;
            double     Phi;
            if ( !frame( u, target, w, Phi ) )
                return NoneLabel;
            double   bestPhi = 1e300;
            uint32_t bestArc = NoneLabel;
            for ( uint32_t t2 = incStart[ a ]; t2 < incStart[ a + 1 ]; t2++ )
            {
                double    ph[ 2 ];
                const int k = crossArc( a, u, w, Phi, incArcs[ t2 ], ph );
                for ( int q = 0; q < k; q++ )
                    if ( ph[ q ] < bestPhi )
                        bestPhi = ph[ q ], bestArc = incArcs[ t2 ];
            }
            return bestArc;
        }
        // an inside point of arc e (see insidePoint) that lies in the face e bounds: the geodesic
        // from it back towards e's cap pole crosses e first; the offset shrinks until it does
        V3<double> insidePointChecked( uint32_t a, uint32_t e, bool & ok ) const
        {
            V3<double> pole;
            double     cosT;
            capOf( a, arcOther( e, a ), pole, cosT );
            V3<double> x { 0, 0, 0 };
            ok = false;
            for ( double delta = 1e-4; delta > 5e-10 && !ok; delta *= 0.1 )
            {
                x  = insidePoint( a, e, delta );
                ok = firstHitTowards( a, x, pole ) == e;
            }
            return x;
        }

        // component of the exposed point c_a + R_a u of sphere a
        uint32_t labelAt( uint32_t a, V3<double> u ) const
        {
            if ( atomLabel[ a ] != MixedLabel )
                return atomLabel[ a ];
            const uint32_t e = firstHit( a, u );
            return e == NoneLabel ? NoneLabel : arcLabel[ e ];
        }
        // a point of arc e just off it on the exposed side of sphere a (unit vector from c_a)
        V3<double> insidePoint( uint32_t a, uint32_t e, double delta ) const
        {
            const uint32_t i = arcU( e ), j = arcV( e );
            V3<double>     ctr, nrm;
            double         rad;
            circle( i, j, ctr, nrm, rad );
            V3<double> x;
            if ( e < Cf )
            {
                const V3<double> ref = std::fabs( nrm.x ) < 0.9 ? V3<double> { 1, 0, 0 } : V3<double> { 0, 1, 0 };
                x                    = ctr + normalize( cross( nrm, ref ) ) * rad;
            }
            else
            {
                const uint32_t * sg = &r.segments[ 4 * size_t( e - Cf ) ];
                const V3<double> A = pos64( sg[ 2 ] ) - ctr, B = pos64( sg[ 3 ] ) - ctr;
                double           ang = std::atan2( dot( cross( A, B ), nrm ), dot( A, B ) );
                if ( ang <= 0 )
                    ang += 2 * M_PI;
                const double h = 0.5 * ang;
                x              = ctr + A * std::cos( h ) + cross( nrm, A ) * std::sin( h );
            }
            V3<double> pole;
            double     cosT;
            capOf( a, arcOther( e, a ), pole, cosT );
            const V3<double> v = normalize( x - c[ a ] );
            const V3<double> t = normalize( v * dot( v, pole ) - pole ); // tangent, away from the cap
            return normalize( v + t * delta );
        }

        // point x of sphere a outside every neighbour's SAS sphere
        bool exposedPoint( uint32_t a, V3<double> x ) const
        {
            for ( uint32_t l = 0; l < r.nbCount[ a ]; l++ )
            {
                const uint32_t j = r.nbIds[ size_t( a ) * r.K + l ];
                if ( len2( x - c[ j ] ) < R[ j ] * R[ j ] )
                    return false;
            }
            return true;
        }
        // atoms bucketed by (y, z) in squares of edge rayCell >= the largest SAS radius: a ray along
        // +x meets only spheres of the 3 x 3 buckets around it
        double                               rayCell = 1, rayOy = 0, rayOz = 0;
        int                                  rayNy = 1, rayNz = 1;
        std::vector<uint32_t>                rayStart, rayAtoms;
        void buildRayIndex()
        {
            const uint32_t n = r.n;
            double         Rm = 0, lo[ 2 ] = { 1e300, 1e300 }, hi[ 2 ] = { -1e300, -1e300 };
            for ( uint32_t a = 0; a < n; a++ )
            {
                Rm = std::max( Rm, R[ a ] );
                lo[ 0 ] = std::min( lo[ 0 ], c[ a ].y ), hi[ 0 ] = std::max( hi[ 0 ], c[ a ].y );
                lo[ 1 ] = std::min( lo[ 1 ], c[ a ].z ), hi[ 1 ] = std::max( hi[ 1 ], c[ a ].z );
            }
            rayCell = std::max( Rm, 0.5 );
            rayOy = lo[ 0 ], rayOz = lo[ 1 ];
            rayNy = n ? int( ( hi[ 0 ] - lo[ 0 ] ) / rayCell ) + 1 : 1, rayNz = n ? int( ( hi[ 1 ] - lo[ 1 ] ) / rayCell ) + 1 : 1;
            rayStart.assign( size_t( rayNy ) * rayNz + 1, 0 );
            auto bucket = [ & ]( uint32_t a ) {
                return size_t( std::min( int( ( c[ a ].z - rayOz ) / rayCell ), rayNz - 1 ) ) * rayNy +
                       size_t( std::min( int( ( c[ a ].y - rayOy ) / rayCell ), rayNy - 1 ) );
            };
            for ( uint32_t a = 0; a < n; a++ )
                rayStart[ bucket( a ) + 1 ]++;
            for ( size_t q = 0; q + 1 < rayStart.size(); q++ )
                rayStart[ q + 1 ] += rayStart[ q ];
            rayAtoms.resize( n );
            std::vector<uint32_t> f( rayStart.begin(), rayStart.end() - 1 );
            for ( uint32_t a = 0; a < n; a++ )
                rayAtoms[ f[ bucket( a ) ]++ ] = a;
        }
        // first SAS sphere (other than `skip`) entered by the ray p + t x^, t > 0: atom and entry x
        uint32_t rayHit( V3<double> p, uint32_t skip, double & hitX ) const
        {
            const int by = int( std::floor( ( p.y - rayOy ) / rayCell ) ), bz = int( std::floor( ( p.z - rayOz ) / rayCell ) );
            uint32_t  best = NoneLabel;
            hitX           = 1e300;
            for ( int z = std::max( bz - 1, 0 ); z <= std::min( bz + 1, rayNz - 1 ); z++ )
                for ( int y = std::max( by - 1, 0 ); y <= std::min( by + 1, rayNy - 1 ); y++ )
                {
                    const size_t bk = size_t( z ) * rayNy + y;
                    for ( uint32_t t = rayStart[ bk ]; t < rayStart[ bk + 1 ]; t++ )
                    {
                        const uint32_t a  = rayAtoms[ t ];
                        const double   dy = c[ a ].y - p.y, dz = c[ a ].z - p.z;
                        const double   h2 = R[ a ] * R[ a ] - dy * dy - dz * dz;
                        if ( a == skip || !( h2 > 0 ) )
                            continue;
                        const double x = c[ a ].x - std::sqrt( h2 );
                        if ( x > p.x && x < hitX )
                            hitX = x, best = a;
                    }
                }
            return best;
        }

        void buildComponents()
        {
            const uint32_t n = r.n;
            Cf = uint32_t( r.fullCircles.size() / 2 ), G = uint32_t( r.segments.size() / 4 ), I = uint32_t( r.concaveIds.size() / 4 );
            const uint32_t E = Cf + G;
            std::vector<uint32_t> par( size_t( E ) + I + n + 1 ); // + the node of the unbounded region
            for ( size_t q = 0; q < par.size(); q++ )
                par[ q ] = uint32_t( q );
            auto find = [ & ]( uint32_t a ) {
                while ( par[ a ] != a )
                    a = par[ a ] = par[ par[ a ] ];
                return a;
            };
            auto unite = [ & ]( uint32_t a, uint32_t b ) {
                a = find( a ), b = find( b );
                if ( a != b )
                    par[ std::max( a, b ) ] = std::min( a, b );
            };
            // a segment joins its end vertices
            for ( uint32_t s = 0; s < G; s++ )
            {
                unite( Cf + s, E + r.segments[ 4 * s + 2 ] );
                unite( Cf + s, E + r.segments[ 4 * s + 3 ] );
            }
            // arcs incident to each atom
            incStart.assign( n + 1, 0 );
            for ( uint32_t e = 0; e < E; e++ )
            {
                incStart[ arcU( e ) + 1 ]++; // both atoms of the arc
                incStart[ arcV( e ) + 1 ]++;
            }
            for ( uint32_t a = 0; a < n; a++ )
                incStart[ a + 1 ] += incStart[ a ];
            incArcs.resize( incStart[ n ] );
            {
                std::vector<uint32_t> f( incStart.begin(), incStart.end() - 1 );
                for ( uint32_t e = 0; e < E; e++ )
                {
                    incArcs[ f[ arcU( e ) ]++ ] = e;
                    incArcs[ f[ arcV( e ) ]++ ] = e;
                }
            }
            // loops of every atom, and the faces they bound
            for ( uint32_t a = 0; a < n; a++ )
            {
                const uint32_t b0 = incStart[ a ], m = incStart[ a + 1 ] - b0;
                if ( m < 2 )
                    continue;
                std::vector<uint32_t> lp( m );
                for ( uint32_t x = 0; x < m; x++ )
                    lp[ x ] = x;
                auto lf = [ & ]( uint32_t x ) {
                    while ( lp[ x ] != x )
                        x = lp[ x ] = lp[ lp[ x ] ];
                    return x;
                };
                auto lu = [ & ]( uint32_t x, uint32_t y ) {
                    x = lf( x ), y = lf( y );
                    if ( x != y )
                        lp[ std::max( x, y ) ] = std::min( x, y );
                };
                // arcs of atom a chain at shared vertices
                for ( uint32_t x = 0; x < m; x++ )
                    for ( uint32_t y = x + 1; y < m; y++ )
                    {
                        const uint32_t ex = incArcs[ b0 + x ], ey = incArcs[ b0 + y ];
                        if ( ex < Cf || ey < Cf )
                            continue;
                        const uint32_t * sx = &r.segments[ 4 * size_t( ex - Cf ) ];
                        const uint32_t * sy = &r.segments[ 4 * size_t( ey - Cf ) ];
                        if ( sx[ 2 ] == sy[ 2 ] || sx[ 2 ] == sy[ 3 ] || sx[ 3 ] == sy[ 2 ] || sx[ 3 ] == sy[ 3 ] )
                            lu( x, y );
                    }
                std::vector<uint32_t> loopOf( m ), reps;
                for ( uint32_t x = 0; x < m; x++ )
                {
                    loopOf[ x ] = lf( x );
                    if ( loopOf[ x ] == x )
                        reps.push_back( x );
                }
                if ( reps.size() < 2 )
                    continue;
                // points just inside each loop's face; two loops bound the same face iff every loop
                // of the atom is crossed an even number of times between them
                std::vector<V3<double>> pts;
                for ( uint32_t r0 : reps )
                {
                    // any arc of the loop bounds the same face: the first whose inside point checks out
                    bool       ok = false;
                    V3<double> pt { 0, 0, 0 };
                    for ( uint32_t x = 0; x < m && !ok; x++ )
                        if ( loopOf[ x ] == r0 )
                            pt = insidePointChecked( a, incArcs[ b0 + x ], ok );
                    pts.push_back( pt );
                }
                for ( size_t p = 0; p < reps.size(); p++ )
                    for ( size_t q = p + 1; q < reps.size(); q++ )
                    {
                        if ( lf( reps[ p ] ) == lf( reps[ q ] ) )
                            continue;
                        // geodesic p -> q, through an intermediate point when nearly antipodal
                        V3<double> path[ 3 ] = { pts[ p ], pts[ q ], pts[ q ] };
                        int        legs      = 1;
                        V3<double> w { 0, 0, 0 };
                        double     Phi = 0;
                        if ( !frame( pts[ p ], pts[ q ], w, Phi ) )
                        {
                            const V3<double> ref = std::fabs( pts[ p ].x ) < 0.9 ? V3<double> { 1, 0, 0 } : V3<double> { 0, 1, 0 };
                            path[ 1 ]            = normalize( cross( pts[ p ], ref ) );
                            legs                 = 2;
                        }
                        std::vector<uint32_t> par2( m, 0 );
                        for ( int l = 0; l < legs; l++ )
                        {
                            frame( path[ l ], path[ l + 1 ], w, Phi );
                            for ( uint32_t x = 0; x < m; x++ )
                            {
                                double ph[ 2 ];
                                par2[ loopOf[ x ] ] += uint32_t( crossArc( a, path[ l ], w, Phi, incArcs[ b0 + x ], ph ) );
                            }
                        }
                        bool even = true;
                        for ( uint32_t x : reps )
                            even = even && par2[ x ] % 2 == 0;
                        if ( even )
                        {
                            lu( reps[ p ], reps[ q ] );
                            unite( incArcs[ b0 + reps[ p ] ], incArcs[ b0 + reps[ q ] ] );
                        }
                    }
                // loops of one face, and the arcs of one loop (joined by vertices already)
                for ( uint32_t x = 0; x < m; x++ )
                    unite( incArcs[ b0 + x ], incArcs[ b0 + loopOf[ x ] ] );
            }
            auto labels = [ & ]() {
            arcLabel.resize( E );
            for ( uint32_t e = 0; e < E; e++ )
                arcLabel[ e ] = find( e );
            vtxLabel.resize( I );
            for ( uint32_t v = 0; v < I; v++ )
                vtxLabel[ v ] = find( E + v );
            atomLabel.assign( n, NoneLabel );
            for ( uint32_t a = 0; a < n; a++ )
            {
                if ( incStart[ a + 1 ] == incStart[ a ] )
                {
                    // no arcs: the exposed region has no boundary, so it is all of the sphere (no
                    // neighbour cuts a cap from it or contains it) or empty (buried)
                    bool buried = false;
                    for ( uint32_t l = 0; l < r.nbCount[ a ] && !buried; l++ )
                    {
                        const uint32_t j = r.nbIds[ size_t( a ) * r.K + l ];
                        const double   d = len( c[ j ] - c[ a ] );
                        buried           = d < R[ a ] + R[ j ] && d + R[ j ] > R[ a ];
                    }
                    atomLabel[ a ] = buried ? NoneLabel : find( E + I + a );
                    continue;
                }
                uint32_t lab = arcLabel[ incArcs[ incStart[ a ] ] ];
                for ( uint32_t t = incStart[ a ] + 1; t < incStart[ a + 1 ]; t++ )
                    if ( arcLabel[ incArcs[ t ] ] != lab )
                        lab = MixedLabel;
                atomLabel[ a ] = lab;
            }
            };
            labels();
            uint32_t best = 0;
            for ( uint32_t a = 1; a < n; a++ )
                if ( double( r.atoms[ 4 * a ] ) + double( r.atoms[ 4 * a + 3 ] ) > double( r.atoms[ 4 * best ] ) + double( r.atoms[ 4 * best + 3 ] ) )
                    best = a;
            (void)best;
            // Regions of the complement: several components may bound one region (the outer
            // surfaces of separate pieces bound the unbounded one; a piece floating in a cavity
            // shares the cavity's region). A ray from an exposed +x pole runs in the region of the
            // pole's face until it enters the boundary of the same region, or escapes (the region is
            // unbounded): the two components are joined, or the pole's joins the unbounded node. The
            // component whose +x extreme point is a pole has such a ray, and along these rays x
            // grows, so every region's components end up joined.
            buildRayIndex();
            const uint32_t extNode = E + I + n;
            for ( uint32_t a = 0; a < n; a++ )
            {
                const V3<double> pole = c[ a ] + V3<double> { R[ a ], 0, 0 };
                if ( !exposedPoint( a, pole ) )
                    continue;
                const uint32_t l0 = labelAt( a, V3<double> { 1, 0, 0 } );
                if ( l0 == NoneLabel )
                    continue;
                double         hx;
                const uint32_t h = rayHit( pole, a, hx );
                uint32_t       l1 = extNode;
                if ( h != NoneLabel )
                {
                    const V3<double> hp { hx, pole.y, pole.z };
                    l1 = labelAt( h, normalize( hp - c[ h ] ) );
                }
                if ( l1 != NoneLabel )
                    unite( l0, l1 );
            }
            labels();
            extComp = find( extNode );
            // exterior status per atom: 0 no exterior boundary, 1 all of it, 2 part of it
            extStatus.assign( n, 0 );
            for ( uint32_t a = 0; a < n; a++ )
            {
                if ( atomLabel[ a ] == MixedLabel )
                {
                    bool any = false, all = true;
                    for ( uint32_t t = incStart[ a ]; t < incStart[ a + 1 ]; t++ )
                    {
                        const bool x = arcLabel[ incArcs[ t ] ] == extComp;
                        any = any || x, all = all && x;
                    }
                    extStatus[ a ] = all ? 1 : any ? 2 : 0;
                }
                else
                    extStatus[ a ] = atomLabel[ a ] == extComp ? 1 : 0;
            }
        }
        std::vector<uint8_t> extStatus;

        void buildGrid( double band )
        {
            const uint32_t n = r.n;
            cell             = std::max( rp + band + rmax + rp, 1.0 );
            double mn[ 3 ] = { 1e300, 1e300, 1e300 }, mx[ 3 ] = { -1e300, -1e300, -1e300 };
            for ( uint32_t a = 0; a < n; a++ )
            {
                mn[ 0 ] = std::min( mn[ 0 ], c[ a ].x ), mx[ 0 ] = std::max( mx[ 0 ], c[ a ].x );
                mn[ 1 ] = std::min( mn[ 1 ], c[ a ].y ), mx[ 1 ] = std::max( mx[ 1 ], c[ a ].y );
                mn[ 2 ] = std::min( mn[ 2 ], c[ a ].z ), mx[ 2 ] = std::max( mx[ 2 ], c[ a ].z );
            }
            ox = mn[ 0 ], oy = mn[ 1 ], oz = mn[ 2 ];
            dx = int( ( mx[ 0 ] - mn[ 0 ] ) / cell ) + 1, dy = int( ( mx[ 1 ] - mn[ 1 ] ) / cell ) + 1, dz = int( ( mx[ 2 ] - mn[ 2 ] ) / cell ) + 1;
            cellStart.assign( size_t( dx ) * dy * dz + 1, 0 );
            std::vector<uint32_t> ca( n );
            for ( uint32_t a = 0; a < n; a++ )
            {
                ca[ a ] = cellOf( c[ a ] );
                cellStart[ ca[ a ] + 1 ]++;
            }
            for ( size_t q = 0; q + 1 < cellStart.size(); q++ )
                cellStart[ q + 1 ] += cellStart[ q ];
            cellAtoms.resize( n );
            std::vector<uint32_t> f( cellStart.begin(), cellStart.end() - 1 );
            for ( uint32_t a = 0; a < n; a++ )
                cellAtoms[ f[ ca[ a ] ]++ ] = a;
        }
        int clampc( double v, double o, int d ) const { return std::min( std::max( int( std::floor( ( v - o ) / cell ) ), 0 ), d - 1 ); }
        uint32_t cellOf( V3<double> p ) const
        {
            return uint32_t( ( size_t( clampc( p.z, oz, dz ) ) * dy + clampc( p.y, oy, dy ) ) * dx + clampc( p.x, ox, dx ) );
        }

        // Vertex positions: stored (as built) or, in dynamic mode, re-solved from the current
        // (possibly perturbed) atoms with the same triplet and side -- topology held fixed.
        bool dynamic = false;
        V3<double> pos64( uint32_t q ) const
        {
            if ( !dynamic )
                return { r.concavePos64[ 3 * q ], r.concavePos64[ 3 * q + 1 ], r.concavePos64[ 3 * q + 2 ] };
            const uint32_t i = uint32_t( r.concaveIds[ 4 * q ] ), j = uint32_t( r.concaveIds[ 4 * q + 1 ] ), k = uint32_t( r.concaveIds[ 4 * q + 2 ] );
            const double   side = r.concavePos[ 4 * q + 3 ] > 0 ? 1.0 : -1.0;
            const Geometry<double> ij = circleGeometry( c[ i ], R[ i ], c[ j ], R[ j ] );
            const Geometry<double> ik = circleGeometry( c[ i ], R[ i ], c[ k ], R[ k ] );
            const V3<double>       u  = ik.normal - dot( ik.normal, ij.normal ) * ij.normal;
            const double           t  = dot( ik.center - ij.center, ik.normal ) / dot( u, ik.normal );
            const V3<double>       x1 = ij.center + t * u;
            const double           cc = std::max( R[ i ] * R[ i ] - len2( x1 - c[ i ] ), 0.0 );
            return x1 + ( side * std::sqrt( cc ) ) * normalize( cross( ij.normal, ik.normal ) );
        }

        // Gradient of w * SDF(q) for the winner `code`, accumulated into g[4 n] (sorted atoms).
        // Inside: nearest boundary point p on spheres m (1 face, 2 arc, 3 vertex); nu solves
        // sum nu_m (p - c_m) = (p - q)/|p - q|; dSDF/dc_m = -nu_m (p - c_m), dSDF/dr_m = -nu_m R_m.
        // Outside: dSDF/dc_a = -(q - c_a)/|q - c_a|, dSDF/dr_a = -1.
        void grad( V3<double> q, int64_t code, double w, double * g ) const
        {
            const int64_t  type = code >> 40;
            const uint32_t idx  = uint32_t( code & 0xffffffffffll );
            if ( type == 5 )
                return;
            if ( type == 0 )
            {
                const V3<double> v = q - c[ idx ];
                const double     d = len( v );
                g[ 4 * idx + 0 ] -= w * v.x / d, g[ 4 * idx + 1 ] -= w * v.y / d, g[ 4 * idx + 2 ] -= w * v.z / d;
                g[ 4 * idx + 3 ] -= w;
                return;
            }
            uint32_t   at[ 3 ];
            int        m = 0;
            V3<double> p;
            if ( type == 1 )
            {
                at[ 0 ] = idx, m = 1;
                const V3<double> v = q - c[ idx ];
                p                  = c[ idx ] + v * ( R[ idx ] / len( v ) );
            }
            else if ( type == 4 )
            {
                at[ 0 ] = uint32_t( r.concaveIds[ 4 * idx ] ), at[ 1 ] = uint32_t( r.concaveIds[ 4 * idx + 1 ] ),
                at[ 2 ] = uint32_t( r.concaveIds[ 4 * idx + 2 ] ), m = 3;
                p = pos64( idx );
            }
            else
            {
                const uint32_t i = type == 2 ? r.fullCircles[ 2 * idx ] : r.segments[ 4 * idx ];
                const uint32_t j = type == 2 ? r.fullCircles[ 2 * idx + 1 ] : r.segments[ 4 * idx + 1 ];
                at[ 0 ] = i, at[ 1 ] = j, m = 2;
                V3<double> ctr, nrm;
                double     rad;
                circle( i, j, ctr, nrm, rad );
                const V3<double> v  = q - ctr;
                const V3<double> wv = v - nrm * dot( v, nrm );
                p                   = ctr + wv * ( rad / len( wv ) );
            }
            const V3<double> e0 = p - q;
            const double     D  = len( e0 );
            if ( !( D > 0 ) )
                return;
            const V3<double> e = e0 * ( 1.0 / D );
            V3<double>       a[ 3 ];
            for ( int u = 0; u < m; u++ )
                a[ u ] = p - c[ at[ u ] ];
            // least-squares nu via the normal equations (Gram matrix), Gaussian elimination
            double G[ 3 ][ 4 ];
            for ( int u = 0; u < m; u++ )
            {
                for ( int v = 0; v < m; v++ )
                    G[ u ][ v ] = dot( a[ u ], a[ v ] );
                G[ u ][ m ] = dot( a[ u ], e );
            }
            for ( int col = 0; col < m; col++ )
            {
                int piv = col;
                for ( int rr = col + 1; rr < m; rr++ )
                    if ( std::fabs( G[ rr ][ col ] ) > std::fabs( G[ piv ][ col ] ) )
                        piv = rr;
                for ( int cc = 0; cc <= m; cc++ )
                    std::swap( G[ col ][ cc ], G[ piv ][ cc ] );
                for ( int rr = 0; rr < m; rr++ )
                    if ( rr != col )
                    {
                        const double f = G[ rr ][ col ] / G[ col ][ col ];
                        for ( int cc = col; cc <= m; cc++ )
                            G[ rr ][ cc ] -= f * G[ col ][ cc ];
                    }
            }
            for ( int u = 0; u < m; u++ )
            {
                const double nu = G[ u ][ m ] / G[ u ][ u ];
                const uint32_t t = at[ u ];
                g[ 4 * t + 0 ] -= w * nu * a[ u ].x, g[ 4 * t + 1 ] -= w * nu * a[ u ].y, g[ 4 * t + 2 ] -= w * nu * a[ u ].z;
                g[ 4 * t + 3 ] -= w * nu * R[ t ];
            }
        }

        // circle (i, j): centre, normal (i -> j), radius
        void circle( uint32_t i, uint32_t j, V3<double> & ctr, V3<double> & nrm, double & rad ) const
        {
            const Geometry<double> g = circleGeometry( c[ i ], R[ i ], c[ j ], R[ j ] );
            ctr = g.center, nrm = g.normal, rad = g.radius;
        }

        double eval( V3<double> q, double band, bool exterior, int64_t & code ) const
        {
            const double D    = rp + band;
            double       g    = 1e300, gExt = 1e300;
            int64_t      gA   = -1;
            double       best = D;
            int64_t      bc   = int64_t( 5 ) << 40;
            const int    cx = clampc( q.x, ox, dx ), cy = clampc( q.y, oy, dy ), cz = clampc( q.z, oz, dz );
            for ( int z = std::max( cz - 1, 0 ); z <= std::min( cz + 1, dz - 1 ); z++ )
                for ( int y = std::max( cy - 1, 0 ); y <= std::min( cy + 1, dy - 1 ); y++ )
                    for ( int x = std::max( cx - 1, 0 ); x <= std::min( cx + 1, dx - 1 ); x++ )
                    {
                        const size_t cc = ( size_t( z ) * dy + y ) * dx + x;
                        for ( uint32_t e = cellStart[ cc ]; e < cellStart[ cc + 1 ]; e++ )
                        {
                            const uint32_t   a    = cellAtoms[ e ];
                            const V3<double> dq   = q - c[ a ];
                            const double     dist = len( dq );
                            const double     sd   = dist - R[ a ];
                            if ( sd < g || ( sd == g && int64_t( a ) < gA ) )
                                g = sd, gA = a;
                            const int st = exterior ? extStatus[ a ] : 1; // 2: check element by element
                            if ( st == 1 )
                                gExt = std::min( gExt, sd );
                            if ( st == 0 )
                                continue;
                            if ( std::fabs( sd ) < best && dist > 0 )
                            {
                                const V3<double> p   = c[ a ] + dq * ( R[ a ] / dist );
                                bool             exp = true;
                                for ( uint32_t l = 0; l < r.nbCount[ a ] && exp; l++ )
                                {
                                    const uint32_t j = r.nbIds[ size_t( a ) * r.K + l ];
                                    exp              = len2( p - c[ j ] ) - R[ j ] * R[ j ] >= 0;
                                }
                                if ( exp && st == 2 )
                                    exp = labelAt( a, dq * ( 1.0 / dist ) ) == extComp;
                                if ( exp )
                                    best = std::fabs( sd ), bc = ( int64_t( 1 ) << 40 ) | a;
                            }
                            for ( uint32_t f = fullStart[ a ]; f < fullStart[ a + 1 ]; f++ )
                            {
                                if ( st == 2 && arcLabel[ f ] != extComp )
                                    continue;
                                V3<double> ctr, nrm;
                                double     rad;
                                circle( r.fullCircles[ 2 * f ], r.fullCircles[ 2 * f + 1 ], ctr, nrm, rad );
                                const V3<double> v  = q - ctr;
                                const double     ax = dot( v, nrm );
                                const V3<double> w  = v - nrm * ax;
                                const double     rho = len( w );
                                const double     d   = std::sqrt( ax * ax + ( rho - rad ) * ( rho - rad ) );
                                if ( d < best )
                                    best = d, bc = ( int64_t( 2 ) << 40 ) | f;
                            }
                            for ( uint32_t s = segStart[ a ]; s < segStart[ a + 1 ]; s++ )
                            {
                                if ( st == 2 && arcLabel[ Cf + s ] != extComp )
                                    continue;
                                const uint32_t * sg = &r.segments[ 4 * size_t( s ) ];
                                V3<double>       ctr, nrm;
                                double           rad;
                                circle( sg[ 0 ], sg[ 1 ], ctr, nrm, rad );
                                const V3<double> v   = q - ctr;
                                const double     ax  = dot( v, nrm );
                                const V3<double> w   = v - nrm * ax;
                                const double     rho = len( w );
                                if ( rho <= 0 )
                                    continue;
                                const double d = std::sqrt( ax * ax + ( rho - rad ) * ( rho - rad ) );
                                if ( d >= best )
                                    continue;
                                const V3<double> A  = pos64( sg[ 2 ] ) - ctr, B = pos64( sg[ 3 ] ) - ctr;
                                auto             an = [ & ]( V3<double> u ) {
                                    double t = std::atan2( dot( cross( A, u ), nrm ), dot( A, u ) );
                                    return t < 0 ? t + 2 * M_PI : t;
                                };
                                if ( an( w ) <= an( B ) )
                                    best = d, bc = ( int64_t( 3 ) << 40 ) | s;
                            }
                            for ( uint32_t t = vtxStart[ a ]; t < vtxStart[ a + 1 ]; t++ )
                            {
                                const uint32_t vq = vtxList[ t ];
                                if ( st == 2 && vtxLabel[ vq ] != extComp )
                                    continue;
                                const double   d  = len( q - pos64( vq ) );
                                if ( d < best )
                                    best = d, bc = ( int64_t( 4 ) << 40 ) | vq;
                            }
                        }
                    }
            double f;
            // no sphere within rp + band: nothing of the surface lies within the band (+band, also
            // for the exterior variant)
            if ( gA < 0 || g >= D )
            {
                code = int64_t( 5 ) << 40;
                if ( exterior )
                {
                    // outside A, farther than rp + band from it: the region holding q is the one
                    // the ray q + t x^ runs in until it enters A (none: unbounded)
                    double         hx;
                    const uint32_t h = rayHit( q, NoneLabel, hx );
                    const uint32_t lh = h == NoneLabel ? extComp : labelAt( h, normalize( V3<double> { hx, q.y, q.z } - c[ h ] ) );
                    if ( lh != NoneLabel && lh != extComp )
                        return -band;
                }
                return band;
            }
            // outside A the point lies in the complement region bounded by the face that holds its
            // projection onto the nearest sphere; ties between components (within 1e-4 A) count as
            // exterior
            bool outsideExt = g >= 0 && ( !exterior || gExt <= g + 1e-4 );
            if ( g >= 0 && !outsideExt && extStatus[ gA ] == 2 )
            {
                const V3<double> dq = q - c[ gA ];
                outsideExt          = labelAt( uint32_t( gA ), dq * ( 1.0 / len( dq ) ) ) == extComp;
            }
            if ( outsideExt )
            {
                f    = g;
                code = gA;
            }
            else
            {
                f    = -best;
                code = bc;
            }
            const double s = rp + f;
            if ( s >= band || s <= -band )
            {
                code = int64_t( 5 ) << 40;
                return s >= band ? band : -band;
            }
            return s;
        }
    };
} // namespace orc

using namespace orc;

extern "C"
{
    // FP64 SDF of a complete-surface oracle build at n query points (xyz doubles); out[n] doubles,
    // codes[n] (optional) winner codes. flags bit 0: exterior variant.
    int oracle_sdf( void * h, const double * pts, uint64_t n, int flags, double band, double * out, int64_t * codes, int threads )
    {
        Result * r = static_cast<Result *>( h );
        if ( r->mode != 0 )
            return 1;
        SdfOracle o( *r );
        o.buildGrid( band );
        parallelFor( uint32_t( n ), threads, [ & ]( uint32_t q ) {
            int64_t code;
            out[ q ] = o.eval( { pts[ 3 * q ], pts[ 3 * q + 1 ], pts[ 3 * q + 2 ] }, band, flags & 1, code );
            if ( codes )
                codes[ q ] = code;
        } );
        return 0;
    }

    // Surface component label per sorted atom and the exterior label (last entry).
    // FP64 analytic gradient of sum_q w_q SDF(q) w.r.t. (x, y, z, r) of every atom, INPUT order (grad[4 n]).
    // Points are split into `threads` contiguous ranges, each summed into its own array; the arrays
    // are then added in range order, so the result depends on the thread count only by rounding.
    int oracle_sdf_grad( void * h, const double * pts, const double * w, uint64_t n, int flags, double band, double * grad, int threads )
    {
        Result * r = static_cast<Result *>( h );
        if ( r->mode != 0 )
            return 1;
        SdfOracle o( *r );
        o.buildGrid( band );
        const int                        T = std::max( 1, std::min( threads, int( ( n + 4095 ) / 4096 ) ) );
        std::vector<std::vector<double>> g( T, std::vector<double>( size_t( r->n ) * 4, 0.0 ) );
        auto                             range = [ & ]( int t ) {
            for ( uint64_t q = n * t / T; q < n * ( t + 1 ) / T; q++ )
            {
                int64_t          code;
                const V3<double> p { pts[ 3 * q ], pts[ 3 * q + 1 ], pts[ 3 * q + 2 ] };
                o.eval( p, band, flags & 1, code );
                o.grad( p, code, w[ q ], g[ t ].data() );
            }
        };
        std::vector<std::thread> pool;
        for ( int t = 1; t < T; t++ )
            pool.emplace_back( range, t );
        range( 0 );
        for ( auto & th : pool )
            th.join();
        for ( uint32_t a = 0; a < r->n; a++ )
            for ( int d = 0; d < 4; d++ )
            {
                double v = 0.0;
                for ( int t = 0; t < T; t++ )
                    v += g[ t ][ 4 * size_t( a ) + d ];
                grad[ 4 * size_t( r->sortedToInitial[ a ] ) + d ] = v;
            }
        return 0;
    }

    // Central finite differences (step hstep, FP64) of sum_q w_q SDF(q) w.r.t. (x, y, z, r) of the
    // listed INPUT atoms, with the element set of the build held fixed and its geometry (circles,
    // vertices) recomputed from the perturbed atoms. stable[k] = 1 when no query near the atom
    // changed its winning element between -h, 0 and +h (the comparison is only meaningful there).
    int oracle_sdf_fd( void * h, const double * pts, const double * w, uint64_t n, int flags, double band, const uint32_t * atomsIn,
                       uint32_t count, double hstep, double * out, uint8_t * stable, int threads )
    {
        Result * r = static_cast<Result *>( h );
        if ( r->mode != 0 )
            return 1;
        std::vector<uint32_t> initToSorted( r->n );
        for ( uint32_t a = 0; a < r->n; a++ )
            initToSorted[ r->sortedToInitial[ a ] ] = a;
        parallelFor( count, threads, [ & ]( uint32_t t ) {
            SdfOracle o( *r );
            o.dynamic = true;
            o.buildGrid( band );
            const uint32_t   a      = initToSorted[ atomsIn[ t ] ];
            const double     reach  = o.rp + band + o.R[ a ] + 2 * ( o.rmax + o.rp ) + 1.0;
            const V3<double> ca     = o.c[ a ];
            std::vector<uint64_t> near;
            for ( uint64_t q = 0; q < n; q++ )
            {
                const V3<double> p { pts[ 3 * q ] - ca.x, pts[ 3 * q + 1 ] - ca.y, pts[ 3 * q + 2 ] - ca.z };
                if ( len2( p ) < reach * reach )
                    near.push_back( q );
            }
            std::vector<int64_t> base( near.size() );
            for ( size_t e = 0; e < near.size(); e++ )
            {
                const uint64_t q = near[ e ];
                o.eval( { pts[ 3 * q ], pts[ 3 * q + 1 ], pts[ 3 * q + 2 ] }, band, flags & 1, base[ e ] );
            }
            bool ok = true;
            for ( int d = 0; d < 4; d++ )
            {
                double L[ 2 ];
                for ( int sgn = 0; sgn < 2; sgn++ )
                {
                    const double step = sgn == 0 ? hstep : -hstep;
                    SdfOracle &  oo   = o;
                    if ( d == 0 ) oo.c[ a ].x = ca.x + step;
                    if ( d == 1 ) oo.c[ a ].y = ca.y + step;
                    if ( d == 2 ) oo.c[ a ].z = ca.z + step;
                    const double R0 = o.R[ a ];
                    if ( d == 3 ) oo.R[ a ] = R0 + step;
                    double acc = 0;
                    for ( size_t e = 0; e < near.size(); e++ )
                    {
                        const uint64_t q = near[ e ];
                        int64_t        code;
                        acc += w[ q ] * oo.eval( { pts[ 3 * q ], pts[ 3 * q + 1 ], pts[ 3 * q + 2 ] }, band, flags & 1, code );
                        if ( code != base[ e ] )
                            ok = false;
                    }
                    L[ sgn ] = acc;
                    oo.c[ a ] = ca;
                    oo.R[ a ] = R0;
                }
                out[ 4 * size_t( t ) + d ] = ( L[ 0 ] - L[ 1 ] ) / ( 2 * hstep );
            }
            stable[ t ] = ok;
        } );
        return 0;
    }

    uint32_t oracle_components( void * h, uint32_t * out )
    {
        Result * r = static_cast<Result *>( h );
        SdfOracle o( *r );
        for ( uint32_t a = 0; a < r->n; a++ )
            out[ a ] = o.atomLabel[ a ];
        return o.extComp;
    }

    // Build the surface of n atoms (float4 xyzr, input order) on the CPU.
    //   mode      0 = complete surface (bcs::Sesdf), 1 = exterior path (bcs::Ssesdf)
    //   precision 32 or 64
    //   track     classify predicates against the epsilon band delta (A^2 / A) and deltaTheta (rad)
    //   threads   host threads for the per-atom stages
    // As oracle_build, with the neighbour lists given per input atom (nbCount[n], nbIds[n * K] of
    // input ids) instead of queried (both null: the grid query), and optionally the sorted order
    // (order[n]: sorted index -> input index; null: the grid's).
    void * oracle_build_nb( const float * xyzr, uint32_t n, float probeRadius, int mode, int precision, uint32_t maxNeighbors,
                            int track, double delta, double deltaTheta, int threads, const uint32_t * nbCount, const uint32_t * nbIds,
                            const uint32_t * order )
    {
        auto * res = new Result();
        res->mode  = mode;
        res->n     = n;
        res->K     = maxNeighbors;
        res->rp    = probeRadius;
        if ( n == 0 )
            return res;
        const AtomF * atoms = reinterpret_cast<const AtomF *>( xyzr );
        if ( precision == 64 )
        {
            Builder<double> b( atoms, n, probeRadius, maxNeighbors, mode, track != 0, delta, deltaTheta, threads, *res );
            b.nbOverrideCount = nbCount, b.nbOverrideIds = nbIds, b.orderOverride = order;
            b.run();
        }
        else
        {
            Builder<float> b( atoms, n, probeRadius, maxNeighbors, mode, false, delta, deltaTheta, threads, *res );
            b.nbOverrideCount = nbCount, b.nbOverrideIds = nbIds, b.orderOverride = order;
            b.run();
        }
        return res;
    }

    void * oracle_build( const float * xyzr, uint32_t n, float probeRadius, int mode, int precision, uint32_t maxNeighbors,
                         int track, double delta, double deltaTheta, int threads )
    {
        return oracle_build_nb( xyzr, n, probeRadius, mode, precision, maxNeighbors, track, delta, deltaTheta, threads, nullptr, nullptr,
                                nullptr );
    }

    void oracle_free( void * h ) { delete static_cast<Result *>( h ); }

    // Named array access: returns the element count and element size (bytes) of array `name`;
    // copies it into out when out is not null.
    uint64_t oracle_get( void * h, const char * name, void * out, uint32_t * elemSize )
    {
        Result *          r = static_cast<Result *>( h );
        const std::string s( name );
        auto              give = [ & ]( const auto & v ) -> uint64_t {
            using E = typename std::decay_t<decltype( v )>::value_type;
            if ( elemSize )
                *elemSize = sizeof( E );
            if ( out && !v.empty() )
                std::memcpy( out, v.data(), v.size() * sizeof( E ) );
            return v.size();
        };
        if ( s == "atoms" ) return give( r->atoms );
        if ( s == "sorted_to_initial" ) return give( r->sortedToInitial );
        if ( s == "nb_count" ) return give( r->nbCount );
        if ( s == "nb_ids" ) return give( r->nbIds );
        if ( s == "status" ) return give( r->status );
        if ( s == "trimmed_to_global" ) return give( r->trimmedToGlobal );
        if ( s == "convex_patches" ) return give( r->convexPatches );
        if ( s == "sectors" ) return give( r->sectors );
        if ( s == "full_circles" ) return give( r->fullCircles );
        if ( s == "concave_ids" ) return give( r->concaveIds );
        if ( s == "concave_pos" ) return give( r->concavePos );
        if ( s == "concave_pos64" ) return give( r->concavePos64 );
        if ( s == "concave_nb" ) return give( r->concaveNb );
        if ( s == "circle_int_nb" ) return give( r->circleIntNb );
        if ( s == "segments" ) return give( r->segments );
        if ( s == "torus" ) return give( r->torus );
        if ( s == "torus_slots" ) return give( r->torusSlots );
        if ( s == "amb_pairs" ) return give( r->ambPairs );
        if ( s == "amb_circle" ) return give( r->ambCircle );
        if ( s == "amb_concave" ) return give( r->ambConcave );
        if ( s == "amb_hasnb" ) return give( r->ambHasNb );
        if ( s == "amb_rejected" ) return give( r->ambRejected );
        if ( s == "amb_seg_circle" ) return give( r->ambSegCircle );
        if ( s == "amb_probe_nb" ) return give( r->ambProbeNb );
        if ( s == "stats" ) return give( r->stats );
        if ( s == "grid" )
        {
            std::vector<float> g { r->grid.ox, r->grid.oy, r->grid.oz, r->grid.cx, r->grid.cy, r->grid.cz, float( r->grid.size ) };
            return give( g );
        }
        if ( s == "counters" )
        {
            std::vector<uint64_t> v { r->Iw, r->anomalyJk, r->overflowNb, r->overflowProbeNb, r->overflowCircleInt };
            return give( v );
        }
        if ( elemSize )
            *elemSize = 0;
        return 0;
    }
}
