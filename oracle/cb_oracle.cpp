// ORACLE — CPU restatement of the reference's Contour-Buildup comparator.
//
// TEST INFRASTRUCTURE ONLY. This file is the checker for the HIP product path (contour_buildup.hip):
// only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it. The product
// library never links it and never falls back to it.
//
// It restates bcs::ContourBuildup::build() (sesdf/src/contour_buildup/contour_buildup.cu:400-660)
// and the kernels it launches (sesdf/extern/ContourBuildup/src/particles_kernel.cu):
//   grid           calcHashD / reorderDataAndFindCellStartD (:67-176), thrust::sort_by_key
//                  (particleSystem.cu:39-44): 16^3 cells over the molecule's box, wrapped hash,
//                  stable sort by hash;
//   neighbours     findNeighborsCBCuda + findNeighborsInCellCBCuda (:1525-1645): cells
//                  [g - c, g + c) per axis (the upper row is not scanned), dist < ri + rj + 2 rp,
//                  first 64 in scan order, small circle (centre vector, radius) per neighbour;
//   covered        removeCoveredSmallCirclesCBCuda (:1647-1769): circles cut off by another
//                  neighbour's plane get radius -1; an atom whose contour vanishes gets no neighbours;
//   arcs           computeArcsCBCuda (:1772-2176): the visible arcs of each small circle, cut by
//                  every other circle in turn (the ten-case interval update), 0 / 2 pi arcs merged,
//                  arc end points with i < j < k written as probes;
//   output         writeProbePositionsCBCuda (:2179-2257): probe positions, spherical-triangle
//                  vectors, torus records of visible circles (i < j), offsets by exclusive scans;
//   singularities  findProbeNeighborsCBCuda (:2259-2379): probes re-hashed on the same grid,
//                  first 32 other probes closer than 2 rp per probe, texture coordinates.
// templated on the scalar type:
//   * T = float  : IEEE FP32 in the reference's operation order (no fast-math emulation);
//   * T = double : the accuracy oracle; with tracking on, every threshold predicate is classified
//                  robust / ambiguous against an epsilon band so that GPU-vs-oracle differences
//                  can be attributed to near-degenerate cases and counted.
// The grid (cell of each atom and probe, stencil extents) is computed in FP32 in both modes: it
// decides which pairs are tested, not what they are.
//
// Parity pins (tests/test_oracle_cb.py): closed forms for two- and three-atom inputs, brute-force
// neighbour lists, and agreement with the independent SES oracle (sesdf_oracle.cpp) on inputs whose
// neighbour lists are complete (probe triples and sides = concave patches, tori = visible circles).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <thread>
#include <type_traits>
#include <vector>

namespace cbo
{
    template<class T>
    struct V3
    {
        T x, y, z;
    };
    template<class T> inline V3<T> operator+( V3<T> a, V3<T> b ) { return { a.x + b.x, a.y + b.y, a.z + b.z }; }
    template<class T> inline V3<T> operator-( V3<T> a, V3<T> b ) { return { a.x - b.x, a.y - b.y, a.z - b.z }; }
    template<class T> inline V3<T> operator*( V3<T> a, T s ) { return { a.x * s, a.y * s, a.z * s }; }
    template<class T> inline V3<T> operator/( V3<T> a, T s ) { return { a.x / s, a.y / s, a.z / s }; }
    template<class T> inline T     dot( V3<T> a, V3<T> b ) { return a.x * b.x + a.y * b.y + a.z * b.z; }
    template<class T> inline T     len( V3<T> a ) { return std::sqrt( dot( a, a ) ); }
    template<class T> inline V3<T> cross( V3<T> a, V3<T> b )
    {
        return { a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x };
    }
    // v * (1 / |v|): the form the HIP kernels use too (correctly rounded sqrt and divide)
    template<class T> inline V3<T> normalize( V3<T> v ) { return v * ( T( 1 ) / std::sqrt( dot( v, v ) ) ); }

    constexpr uint32_t K       = 64; // atomNeighborCount (contour_buildup.cu:57)
    constexpr uint32_t ProbeK  = 32; // probeNeighborCount (contour_buildup.cu:57)
    constexpr uint32_t Empty   = 0xffffffffu;
    constexpr uint32_t ArcCap  = 64; // start[64] ... (particles_kernel.cu:1810-1815)
    constexpr uint32_t TmpCap  = 16; // tmpStart[16] ... (:1822-1826)

    struct Params
    {
        float    origin[ 3 ];
        float    cell[ 3 ];
        uint32_t dim;     // cells per axis (power of two)
        float    rp;
        uint32_t texSize; // GL_MAX_TEXTURE_SIZE of the reference's context
    };

    // calcGridPos / calcGridHash (particles_kernel.cu:64-84), FP32
    inline void grid_pos( const Params & p, float x, float y, float z, int g[ 3 ] )
    {
        g[ 0 ] = int( std::floor( ( x - p.origin[ 0 ] ) / p.cell[ 0 ] ) );
        g[ 1 ] = int( std::floor( ( y - p.origin[ 1 ] ) / p.cell[ 1 ] ) );
        g[ 2 ] = int( std::floor( ( z - p.origin[ 2 ] ) / p.cell[ 2 ] ) );
    }
    inline uint32_t grid_hash( const Params & p, int gx, int gy, int gz )
    {
        const int m = int( p.dim ) - 1;
        return ( uint32_t( gz & m ) * p.dim + uint32_t( gy & m ) ) * p.dim + uint32_t( gx & m );
    }

    struct Grid
    {
        std::vector<uint32_t> sortedToInitial, cellStart, cellEnd;
    };
    // stable sort of points by cell hash, cell ranges (thrust::sort_by_key + reorderDataAndFindCellStartD)
    inline Grid build_grid( const Params & p, const float * xyzw, uint32_t n )
    {
        Grid                  g;
        std::vector<uint32_t> hash( n );
        for ( uint32_t i = 0; i < n; i++ )
        {
            int c[ 3 ];
            grid_pos( p, xyzw[ 4 * i ], xyzw[ 4 * i + 1 ], xyzw[ 4 * i + 2 ], c );
            hash[ i ] = grid_hash( p, c[ 0 ], c[ 1 ], c[ 2 ] );
        }
        g.sortedToInitial.resize( n );
        for ( uint32_t i = 0; i < n; i++ )
            g.sortedToInitial[ i ] = i;
        std::stable_sort( g.sortedToInitial.begin(), g.sortedToInitial.end(),
                          [ & ]( uint32_t a, uint32_t b ) { return hash[ a ] < hash[ b ]; } );
        const uint32_t cells = p.dim * p.dim * p.dim;
        g.cellStart.assign( cells, Empty );
        g.cellEnd.assign( cells, 0 );
        for ( uint32_t s = 0; s < n; s++ )
        {
            const uint32_t h = hash[ g.sortedToInitial[ s ] ];
            if ( s == 0 || h != hash[ g.sortedToInitial[ s - 1 ] ] )
            {
                g.cellStart[ h ] = s;
                if ( s > 0 )
                    g.cellEnd[ hash[ g.sortedToInitial[ s - 1 ] ] ] = s;
            }
            if ( s == n - 1 )
                g.cellEnd[ h ] = s + 1;
        }
        return g;
    }

    // Visits the stencil cells [g - c, g + c) (clamped to the grid) in z, y, x order.
    template<class F>
    inline void for_stencil( const Params & p, const int g[ 3 ], float range, F && visit )
    {
        int c[ 3 ];
        for ( int a = 0; a < 3; a++ )
            c[ a ] = int( std::ceil( range / p.cell[ a ] ) );
        const int dim = int( p.dim );
        for ( int z = std::max( g[ 2 ] - c[ 2 ], 0 ); z < std::min( g[ 2 ] + c[ 2 ], dim ); z++ )
            for ( int y = std::max( g[ 1 ] - c[ 1 ], 0 ); y < std::min( g[ 1 ] + c[ 1 ], dim ); y++ )
                for ( int x = std::max( g[ 0 ] - c[ 0 ], 0 ); x < std::min( g[ 0 ] + c[ 0 ], dim ); x++ )
                    if ( !visit( grid_hash( p, x, y, z ) ) )
                        return;
    }

    struct Result
    {
        uint32_t              n = 0;
        Params                prm {};
        int                   precision = 32;
        std::vector<float>    atoms;            // 4 N, grid-sorted
        std::vector<uint32_t> sortedToInitial;  // N
        std::vector<uint32_t> nbCount0;         // N: found neighbours
        std::vector<uint32_t> nbCount;          // N: after the covered-circle pass (0: no contour)
        std::vector<uint32_t> nbIds;            // N K (sorted atom indices)
        std::vector<uint8_t>  removed;          // N K: circle cut off (radius -1)
        std::vector<uint8_t>  arcCount;         // N K: probes written from the circle
        std::vector<uint8_t>  visible;          // N K: torus written (i < j, at least one arc)
        std::vector<float>    probePos, tria1, tria2, tria3; // 4 P each
        std::vector<uint32_t> probeIds;         // 4 P: (i, j, k, 0) sorted atom indices
        std::vector<double>   probePos64;       // 3 P (FP64 builds)
        std::vector<float>    torusPos, torusAxis, torusVs; // 4 T each
        std::vector<uint32_t> torusIds;         // 2 T
        std::vector<uint32_t> probeSorted;      // P: sorted probe -> probe index
        std::vector<float>    texCoord;         // 3 P (probe order)
        std::vector<float>    singTex;          // 3 * 32 P (sorted probe order; unused slots 0)
        std::vector<uint8_t>  ambAtom;          // N: neighbour or no-contour decision in the band
        std::vector<uint8_t>  ambCover;         // N K: covered-circle decision in the band
        std::vector<uint8_t>  ambCircle;        // N K: arc decision in the band
        std::vector<uint8_t>  ambProbe;         // P: a probe pair in the band of 2 rp
        std::vector<uint64_t> stats;            // see StatNames
    };
    enum Stat
    {
        StatCandidates,   // atom pairs tested for neighbourhood
        StatCoverTests,   // (circle, k) pairs of the covered-circle pass
        StatArcTests,     // (circle, k) pairs of the arc pass reaching the interval update
        StatArcUpdates,   // (arc, k) interval updates
        StatProbePairs,   // probe pairs tested for singularities
        StatArcPeak,      // largest arc list (arcs + new arcs of one cut) seen
        StatNbOverflow,   // atoms whose neighbour list hit 64
        StatArcOverflow,  // circles whose arc list exceeded the reference's 64 + 16 slots
        StatCount
    };

    template<class T>
    struct Builder
    {
        const Params & p;
        Result &       R;
        bool           track;
        T              delta;  // A, distance-like comparisons
        T              dtheta; // rad, angle comparisons
        std::vector<V3<T>> vc;  // N K: small-circle centre vectors
        std::vector<T>     wc;  // N K: small-circle radii

        static T twoPi() { return std::is_same<T, float>::value ? T( 6.28318530718f ) : T( 6.283185307179586 ); }

        // Bound on the FP32 error of the point h where two circle planes meet: the solve amplifies
        // rounding (about 6e-8 per operation) by the conditioning |vj|^2 |vk|^2 / denom.
        static T line_error( V3<T> vj, V3<T> vk, T denom, V3<T> h )
        {
            const T cond = dot( vj, vj ) * dot( vk, vk ) / std::max( std::fabs( denom ), T( 1e-30 ) );
            return T( 1e-6 ) * std::max( T( 1 ), cond ) * ( len( h ) + len( vj ) + len( vk ) );
        }

        V3<T> at( uint32_t s ) const { return { T( R.atoms[ 4 * s ] ), T( R.atoms[ 4 * s + 1 ] ), T( R.atoms[ 4 * s + 2 ] ) }; }
        T     rad( uint32_t s ) const { return T( R.atoms[ 4 * s + 3 ] ); }

        // ---- neighbours and small circles (findNeighborsCBCuda) ----
        void neighbours( const Grid & g, uint32_t i, uint64_t & cand )
        {
            const float * a = &R.atoms[ 4 * i ];
            int           gp[ 3 ];
            grid_pos( p, a[ 0 ], a[ 1 ], a[ 2 ], gp );
            const float range = a[ 3 ] + 3.0f + 2.0f * p.rp;
            const V3<T> pos   = at( i );
            const T     rpT   = T( p.rp );
            uint32_t    count = 0;
            bool        full  = false;
            for_stencil( p, gp, range, [ & ]( uint32_t h ) {
                const uint32_t s = g.cellStart[ h ];
                if ( s == Empty )
                    return true;
                for ( uint32_t j = s; j < g.cellEnd[ h ]; j++ )
                {
                    if ( j == i )
                        continue;
                    cand++;
                    const V3<T> rel  = at( j ) - pos;
                    const T     dist = len( rel );
                    const T     nd   = rad( i ) + rad( j ) + T( 2 ) * rpT;
                    if ( track && std::fabs( dist - nd ) <= delta )
                        R.ambAtom[ i ] = 1;
                    if ( !( dist < nd ) )
                        continue;
                    if ( count >= K )
                    {
                        full = true;
                        return false;
                    }
                    R.nbIds[ size_t( i ) * K + count ] = j;
                    const T Ri = rad( i ) + rpT, Rj = rad( j ) + rpT;
                    T       r  = Ri * Ri + dist * dist - Rj * Rj;
                    r          = r / ( T( 2 ) * dist * dist );
                    const V3<T> v = rel * r;
                    vc[ size_t( i ) * K + count ] = v;
                    wc[ size_t( i ) * K + count ] = std::sqrt( Ri * Ri - dot( v, v ) );
                    count++;
                }
                return true;
            } );
            R.nbCount0[ i ] = count;
            if ( full )
                R.stats[ StatNbOverflow ]++;
        }

        // ---- covered circles (removeCoveredSmallCirclesCBCuda) ----
        // Returns true when atom i has no contour.
        bool covered( uint32_t i, uint64_t & tests )
        {
            const uint32_t n  = R.nbCount0[ i ];
            const V3<T>    pi = at( i );
            const T        Rr = rad( i ) + T( p.rp );
            bool           noContour = false;
            for ( uint32_t jj = 0; jj < n; jj++ )
            {
                const size_t sj = size_t( i ) * K + jj;
                const V3<T>  vj = vc[ sj ], pj = at( R.nbIds[ sj ] );
                bool         addJ = true;
                for ( uint32_t kk = 0; kk < n; kk++ )
                {
                    if ( kk == jj )
                        continue;
                    tests++;
                    const size_t sk    = size_t( i ) * K + kk;
                    const V3<T>  vk    = vc[ sk ], pk = at( R.nbIds[ sk ] );
                    const T      vjvk  = dot( vj, vk );
                    const T      denom = dot( vj, vj ) * dot( vk, vk ) - vjvk * vjvk;
                    const V3<T>  h     = vj * ( dot( vj, vj - vk ) * dot( vk, vk ) ) / denom
                                    + vk * ( dot( vk - vj, vk ) * dot( vj, vj ) ) / denom;
                    const V3<T> nj = normalize( pi - pj ), nk = normalize( pi - pk ), q = vk - vj;
                    const T     njnk = dot( nj, nk );
                    if ( njnk == T( 1 ) )
                    {
                        if ( dot( nj, q ) > T( 0 ) )
                            addJ = false;
                        if ( track )
                            R.ambCover[ sj ] = 1;
                        continue;
                    }
                    const T     lh = len( h );
                    const V3<T> mj = vj - h, mk = vk - h;
                    const T     mm = dot( mj, mk ), nq = dot( nj, q );
                    if ( track )
                    {
                        // three-valued outcome of this test: 1 true, 0 false, -1 inside the band
                        auto tv = []( bool v, bool amb ) { return amb ? -1 : int( v ); };
                        auto and3 = []( int a, int b ) { return ( a == 0 || b == 0 ) ? 0 : ( a == 1 && b == 1 ) ? 1 : -1; };
                        const T    eh         = line_error( vj, vk, denom, h );
                        const bool degenerate = T( 1 ) - njnk <= T( 1e-6 );
                        const int  far  = tv( lh > Rr, std::fabs( lh - Rr ) <= delta + eh );
                        const int  same = tv( njnk > T( 0 ), std::fabs( njnk ) <= T( 1e-6 ) );
                        const int  opp  = same < 0 ? -1 : 1 - same;
                        const T    lmj = len( mj ), lmk = len( mk );
                        const int  mPos = tv( mm > T( 0 ), std::fabs( mm ) <= T( 1e-5 ) * lmj * lmk + eh * ( lmj + lmk + eh ) );
                        const int  qPos = tv( nq > T( 0 ), std::fabs( nq ) <= delta );
                        const int  qNeg = qPos < 0 ? -1 : 1 - qPos;
                        if ( degenerate || and3( and3( far, same ), and3( mPos, qPos ) ) < 0 )
                            R.ambCover[ sj ] = 1;
                        if ( degenerate || and3( and3( far, opp ), and3( mPos, qNeg ) ) < 0 )
                            R.ambAtom[ i ] = 1;
                    }
                    if ( !( lh > Rr ) )
                        continue;
                    if ( njnk > T( 0 ) )
                    {
                        if ( mm > T( 0 ) && nq > T( 0 ) )
                            addJ = false;
                    }
                    else if ( mm > T( 0 ) && nq < T( 0 ) )
                        noContour = true;
                }
                R.removed[ sj ] = !addJ;
            }
            return noContour;
        }

        // ---- arcs (computeArcsCBCuda) ----
        struct Arc
        {
            T        s, e;
            uint32_t sk, ek;
            bool     valid;
        };
        struct ProbeOut
        {
            V3<T>    x;
            uint32_t k;
        };

        void arcs( uint32_t i, uint32_t jj, std::vector<ProbeOut> & out, uint64_t & tests, uint64_t & updates, uint64_t & peak,
                   uint64_t & overflow )
        {
            const size_t sj  = size_t( i ) * K + jj;
            const uint32_t n = R.nbCount[ i ];
            if ( wc[ sj ] < T( 0 ) || R.removed[ sj ] )
                return;
            const T        TP = twoPi();
            const V3<T>    pi = at( i );
            const T        Rr = rad( i ) + T( p.rp );
            const uint32_t j  = R.nbIds[ sj ];
            const V3<T>    vj = vc[ sj ];
            const T        w  = wc[ sj ];
            // angle band: the FP32 angle error grows as the circle shrinks
            const T        band = dtheta + delta / std::max( w, T( 1e-3 ) );
            bool           amb  = false;
            auto           near = [ & ]( T a, T b ) {
                if ( track && std::fabs( a - b ) <= band )
                    amb = true;
            };
            // fmod(x, 2 pi) for x >= 0 jumps at multiples of 2 pi
            auto wrapped = [ & ]( T x ) {
                // exactly 2 pi is the untouched end of the initial arc (a constant in every build)
                if ( track && x > band && x != TP )
                {
                    const T m = std::fmod( x, TP );
                    if ( m <= band || TP - m <= band )
                        amb = true;
                }
                return std::fmod( x, TP );
            };

            std::vector<Arc> arcsV;
            arcsV.reserve( 8 );
            arcsV.push_back( { T( 0 ), TP, 0u, 0u, true } );
            V3<T> xAxis = cross( vj, V3<T> { T( 0 ), T( 1 ), T( 0 ) } );
            if ( dot( xAxis, xAxis ) == T( 0 ) )
                xAxis = cross( vj, V3<T> { T( 1 ), T( 0 ), T( 0 ) } );
            xAxis             = normalize( xAxis );
            const V3<T> yAxis = normalize( cross( xAxis, vj ) );

            for ( uint32_t kk = 0; kk < n; kk++ )
            {
                if ( kk == jj )
                    continue;
                const size_t sk = size_t( i ) * K + kk;
                if ( wc[ sk ] < T( 0 ) || R.removed[ sk ] )
                    continue;
                const uint32_t k     = R.nbIds[ sk ];
                const V3<T>    vk    = vc[ sk ], pk = at( k );
                const T        vjvk  = dot( vj, vk );
                const T        denom = dot( vj, vj ) * dot( vk, vk ) - vjvk * vjvk;
                const V3<T>    h     = vj * ( dot( vj, vj - vk ) * dot( vk, vk ) ) / denom
                                + vk * ( dot( vk - vj, vk ) * dot( vj, vj ) ) / denom;
                const T lh = len( h );
                if ( track && std::fabs( lh - Rr ) <= delta + line_error( vj, vk, denom, h ) )
                    amb = true;
                if ( lh > Rr )
                    continue;
                tests++;
                const V3<T> cr   = cross( vk, vj );
                const T     root = std::sqrt( ( Rr * Rr - dot( h, h ) ) / dot( cr, cr ) );
                V3<T>       x1 = h + cr * root, x2 = h - cr * root;
                const T     side = dot( vk, pk - pi );
                if ( track && std::fabs( side ) <= delta * len( vk ) )
                    amb = true;
                if ( side < T( 0 ) )
                    std::swap( x1, x2 );
                T a1 = std::atan2( dot( x1 - vj, yAxis ), dot( x1 - vj, xAxis ) );
                T a2 = std::atan2( dot( x2 - vj, yAxis ), dot( x2 - vj, xAxis ) );
                if ( a1 > TP )
                {
                    a1 = std::fmod( a1, TP );
                    a2 = std::fmod( a2, TP );
                }
                near( a2, a1 );
                if ( a2 < a1 )
                    a2 += TP;
                near( a1, T( 0 ) );
                if ( a1 < T( 0 ) )
                {
                    a1 += TP;
                    a2 += TP;
                }

                const size_t before = arcsV.size();
                for ( size_t a = 0; a < before; a++ )
                {
                    if ( !arcsV[ a ].valid )
                        continue;
                    updates++;
                    const T        s = arcsV[ a ].s, e = arcsV[ a ].e;
                    // the reference holds the k indices of the arc being split in floats
                    const uint32_t skF = uint32_t( float( arcsV[ a ].sk ) ), ekF = uint32_t( float( arcsV[ a ].ek ) );
                    Arc &          A   = arcsV[ a ];
                    auto           fixEnd = [ & ]( T & en, T st ) {
                        near( en, st );
                        if ( en < st )
                            en += TP;
                    };
                    near( a1, s );
                    if ( a1 < s )
                    {
                        const T d1 = s - a1, span = a2 - a1, reach = d1 + ( e - s );
                        near( d1, span );
                        if ( d1 > span )
                        {
                            near( reach, TP );
                            if ( reach > TP )
                            {
                                near( reach, TP + a2 - a1 );
                                if ( reach < TP + a2 - a1 )
                                {
                                    A.s = a1, A.sk = k, A.e = wrapped( e );
                                    fixEnd( A.e, A.s );
                                }
                                else
                                {
                                    A.s = a1, A.sk = k, A.e = wrapped( a2 ), A.ek = k;
                                    fixEnd( A.e, A.s );
                                }
                            }
                            else
                                A.valid = false;
                        }
                        else
                        {
                            near( reach, span );
                            if ( reach > span )
                            {
                                A.e = wrapped( a2 ), A.ek = k;
                                fixEnd( A.e, A.s );
                                near( reach, TP );
                                if ( reach > TP )
                                {
                                    Arc t { a1, wrapped( e ), k, ekF, true };
                                    fixEnd( t.e, t.s );
                                    arcsV.push_back( t );
                                }
                            }
                        }
                    }
                    else
                    {
                        const T d1 = a1 - s, reach = d1 + ( a2 - a1 ), span = e - s;
                        near( d1, span );
                        if ( d1 > span )
                        {
                            near( reach, TP );
                            if ( reach > TP )
                            {
                                near( reach, TP + e - s );
                                if ( reach < TP + e - s )
                                {
                                    A.e = wrapped( a2 ), A.ek = k;
                                    fixEnd( A.e, A.s );
                                }
                            }
                            else
                                A.valid = false;
                        }
                        else
                        {
                            near( reach, span );
                            if ( reach > span )
                            {
                                A.s = a1, A.sk = k, A.e = wrapped( A.e );
                                fixEnd( A.e, A.s );
                                near( reach, TP );
                                if ( reach > TP )
                                {
                                    Arc t { s, wrapped( a2 ), skF, k, true };
                                    fixEnd( t.e, t.s );
                                    arcsV.push_back( t );
                                }
                            }
                            else
                            {
                                A.s = a1, A.sk = k, A.e = wrapped( a2 ), A.ek = k;
                                fixEnd( A.e, A.s );
                            }
                        }
                    }
                }
                peak = std::max<uint64_t>( peak, arcsV.size() );
                if ( arcsV.size() > ArcCap || arcsV.size() - before > TmpCap )
                    overflow++;
                size_t o = 0;
                for ( size_t a = 0; a < arcsV.size(); a++ )
                    if ( arcsV[ a ].valid )
                        arcsV[ o++ ] = arcsV[ a ];
                arcsV.resize( o );
            }

            // merge an arc starting at 0 with one ending at 2 pi
            int idx0 = -1, idx2 = -1;
            for ( size_t a = 0; a < arcsV.size(); a++ )
            {
                const T s = arcsV[ a ].s, e = arcsV[ a ].e;
                if ( s != T( 0 ) )
                    near( s, T( 1e-5 ) );
                if ( s < T( 1e-5 ) )
                    idx0 = int( a );
                else
                {
                    if ( e != TP )
                        near( std::fabs( e - TP ), T( 1e-4 ) );
                    if ( std::fabs( e - TP ) < T( 1e-4 ) )
                        idx2 = int( a );
                }
            }
            if ( idx0 >= 0 && idx2 >= 0 )
            {
                Arc & A = arcsV[ idx0 ];
                A.s     = arcsV[ idx2 ].s;
                A.sk    = arcsV[ idx2 ].sk;
                A.e     = wrapped( A.e );
                near( A.e, A.s );
                if ( A.e < A.s )
                    A.e += TP;
                arcsV.erase( arcsV.begin() + idx2 );
            }

            uint32_t written = 0;
            for ( const Arc & A : arcsV )
            {
                if ( i < j )
                {
                    if ( j < A.sk )
                    {
                        out.push_back( { pi + vj + ( xAxis * std::cos( A.s ) + yAxis * std::sin( A.s ) ) * w, A.sk } );
                        written++;
                    }
                    if ( j < A.ek )
                    {
                        out.push_back( { pi + vj + ( xAxis * std::cos( A.e ) + yAxis * std::sin( A.e ) ) * w, A.ek } );
                        written++;
                    }
                }
            }
            R.arcCount[ sj ] = uint8_t( std::min<uint32_t>( written, 255 ) );
            R.visible[ sj ]  = ( i < j && !arcsV.empty() ) ? 1 : 0;
            if ( amb )
                R.ambCircle[ sj ] = 1;
        }

        // Torus record of a visible circle (writeProbePositionsCBCuda, :2237-2256)
        void torus( uint32_t i, uint32_t jj, float * pos, float * axis, float * vs )
        {
            const size_t sj = size_t( i ) * K + jj;
            const V3<T>  ai = at( i ), aj = at( R.nbIds[ sj ] );
            const V3<T>  sc = vc[ sj ];
            const T      w  = wc[ sj ];
            const V3<T>  ta = normalize( sc );
            const V3<T>  tc = sc + ai;
            const V3<T>  ortho = normalize( cross( ta, V3<T> { T( 0 ), T( 0 ), T( 1 ) } ) );
            const V3<T>  P     = tc + ortho * w;
            const V3<T>  X     = normalize( P - ai ) * rad( i );
            V3<T>        C     = ( aj - ai ) * ( len( P - ai ) / ( len( P - aj ) + len( P - ai ) ) );
            const T      dist  = len( X - C );
            C                  = ( C + ai ) - tc;
            const float v[ 12 ] = { float( tc.x ), float( tc.y ), float( tc.z ), float( w ),
                                    float( ta.x ), float( ta.y ), float( ta.z ), p.rp,
                                    float( C.x ),  float( C.y ),  float( C.z ),  float( dist ) };
            std::memcpy( pos, v, 16 );
            std::memcpy( axis, v + 4, 16 );
            std::memcpy( vs, v + 8, 16 );
        }
    };

    // Singularity texture (findProbeNeighborsCBCuda) over the probe array, FP32.
    inline void singularities( Result & R, const Params & p, bool track, double delta )
    {
        const uint32_t P = uint32_t( R.probePos.size() / 4 );
        R.probeSorted.clear();
        R.texCoord.assign( size_t( P ) * 3, 0.f );
        R.singTex.assign( size_t( P ) * ProbeK * 3, 0.f );
        R.ambProbe.assign( P, 0 );
        if ( P == 0 )
            return;
        Grid g        = build_grid( p, R.probePos.data(), P );
        R.probeSorted = g.sortedToInitial;
        const float nd = p.rp + p.rp;
        uint64_t    pairs = 0;
        for ( uint32_t s = 0; s < P; s++ )
        {
            const float * a = &R.probePos[ 4 * size_t( g.sortedToInitial[ s ] ) ];
            int           gp[ 3 ];
            grid_pos( p, a[ 0 ], a[ 1 ], a[ 2 ], gp );
            uint32_t count = 0;
            for_stencil( p, gp, nd, [ & ]( uint32_t h ) {
                const uint32_t st = g.cellStart[ h ];
                if ( st == Empty )
                    return true;
                for ( uint32_t t = st; t < g.cellEnd[ h ]; t++ )
                {
                    if ( t == s )
                        continue;
                    pairs++;
                    const float * b  = &R.probePos[ 4 * size_t( g.sortedToInitial[ t ] ) ];
                    const float   dx = b[ 0 ] - a[ 0 ], dy = b[ 1 ] - a[ 1 ], dz = b[ 2 ] - a[ 2 ];
                    const float   d  = std::sqrt( dx * dx + dy * dy + dz * dz );
                    if ( track && std::fabs( double( d ) - double( nd ) ) <= delta )
                        R.ambProbe[ g.sortedToInitial[ s ] ] = 1;
                    if ( !( d < nd ) )
                        continue;
                    if ( count >= ProbeK )
                        return false;
                    float * o = &R.singTex[ ( size_t( s ) * ProbeK + count ) * 3 ];
                    o[ 0 ] = b[ 0 ], o[ 1 ] = b[ 1 ], o[ 2 ] = b[ 2 ];
                    count++;
                }
                return true;
            } );
            float * tc = &R.texCoord[ size_t( g.sortedToInitial[ s ] ) * 3 ];
            tc[ 0 ] = float( count ), tc[ 1 ] = float( s / p.texSize ), tc[ 2 ] = float( s % p.texSize );
        }
        R.stats[ StatProbePairs ] += pairs;
    }

    template<class T>
    Result * build( const float * xyzr, uint32_t n, const Params & p, bool track, double delta, double dtheta, int threads )
    {
        Result * Rp = new Result();
        Result & R  = *Rp;
        R.n = n, R.prm = p, R.precision = std::is_same<T, float>::value ? 32 : 64;
        R.stats.assign( StatCount, 0 );
        Grid g            = build_grid( p, xyzr, n );
        R.sortedToInitial = g.sortedToInitial;
        R.atoms.resize( size_t( n ) * 4 );
        for ( uint32_t s = 0; s < n; s++ )
            std::memcpy( &R.atoms[ 4 * size_t( s ) ], xyzr + 4 * size_t( g.sortedToInitial[ s ] ), 16 );
        R.nbCount0.assign( n, 0 );
        R.nbCount.assign( n, 0 );
        R.nbIds.assign( size_t( n ) * K, 0 );
        R.removed.assign( size_t( n ) * K, 0 );
        R.arcCount.assign( size_t( n ) * K, 0 );
        R.visible.assign( size_t( n ) * K, 0 );
        R.ambAtom.assign( n, 0 );
        R.ambCover.assign( size_t( n ) * K, 0 );
        R.ambCircle.assign( size_t( n ) * K, 0 );

        Builder<T> B { p, R, track, T( delta ), T( dtheta ), {}, {} };
        B.vc.assign( size_t( n ) * K, V3<T> { 0, 0, 0 } );
        B.wc.assign( size_t( n ) * K, T( 0 ) );

        // per-atom work, split over threads in contiguous ranges; results gathered in atom order
        const int nt = std::max( 1, std::min( threads, int( n / 64 ) + 1 ) );
        std::vector<std::vector<typename Builder<T>::ProbeOut>> probes( n );
        std::vector<uint64_t> st( size_t( nt ) * StatCount, 0 );
        auto run = [ & ]( auto && fn ) {
            std::vector<std::thread> pool;
            for ( int t = 0; t < nt; t++ )
                pool.emplace_back( [ &, t ] {
                    const uint32_t lo = uint32_t( uint64_t( n ) * t / nt ), hi = uint32_t( uint64_t( n ) * ( t + 1 ) / nt );
                    for ( uint32_t i = lo; i < hi; i++ )
                        fn( i, &st[ size_t( t ) * StatCount ] );
                } );
            for ( auto & th : pool )
                th.join();
        };
        run( [ & ]( uint32_t i, uint64_t * s ) { B.neighbours( g, i, s[ StatCandidates ] ); } );
        run( [ & ]( uint32_t i, uint64_t * s ) {
            const bool none = B.covered( i, s[ StatCoverTests ] );
            R.nbCount[ i ]  = none ? 0 : R.nbCount0[ i ];
        } );
        run( [ & ]( uint32_t i, uint64_t * s ) {
            uint64_t peak = 0;
            for ( uint32_t jj = 0; jj < R.nbCount[ i ]; jj++ )
                B.arcs( i, jj, probes[ i ], s[ StatArcTests ], s[ StatArcUpdates ], peak, s[ StatArcOverflow ] );
            s[ StatArcPeak ] = std::max( s[ StatArcPeak ], peak );
        } );
        for ( int t = 0; t < nt; t++ )
            for ( int c = 0; c < StatCount; c++ )
                R.stats[ c ] = c == StatArcPeak ? std::max( R.stats[ c ], st[ size_t( t ) * StatCount + c ] )
                                                : R.stats[ c ] + st[ size_t( t ) * StatCount + c ];

        // probes and tori in (atom, neighbour slot, arc) order (the reference's exclusive scans)
        for ( uint32_t i = 0; i < n; i++ )
        {
            size_t pOff = 0;
            for ( uint32_t jj = 0; jj < R.nbCount[ i ]; jj++ )
            {
                const size_t   sj = size_t( i ) * K + jj;
                const uint32_t j  = R.nbIds[ sj ];
                for ( uint32_t c = 0; c < R.arcCount[ sj ]; c++ )
                {
                    const auto & pr = probes[ i ][ pOff++ ];
                    const float  x = float( pr.x.x ), y = float( pr.x.y ), z = float( pr.x.z );
                    const float * ai = &R.atoms[ 4 * size_t( i ) ], *aj = &R.atoms[ 4 * size_t( j ) ], *ak = &R.atoms[ 4 * size_t( pr.k ) ];
                    const float v[ 16 ] = { x, y, z, p.rp,
                                            ai[ 0 ] - x, ai[ 1 ] - y, ai[ 2 ] - z, 1.f,
                                            aj[ 0 ] - x, aj[ 1 ] - y, aj[ 2 ] - z, 1.f,
                                            ak[ 0 ] - x, ak[ 1 ] - y, ak[ 2 ] - z, p.rp * p.rp };
                    R.probePos.insert( R.probePos.end(), v, v + 4 );
                    R.tria1.insert( R.tria1.end(), v + 4, v + 8 );
                    R.tria2.insert( R.tria2.end(), v + 8, v + 12 );
                    R.tria3.insert( R.tria3.end(), v + 12, v + 16 );
                    const uint32_t ids[ 4 ] = { i, j, pr.k, 0u };
                    R.probeIds.insert( R.probeIds.end(), ids, ids + 4 );
                    R.probePos64.push_back( double( pr.x.x ) );
                    R.probePos64.push_back( double( pr.x.y ) );
                    R.probePos64.push_back( double( pr.x.z ) );
                }
                if ( R.visible[ sj ] )
                {
                    float tp[ 4 ], ta[ 4 ], tv[ 4 ];
                    B.torus( i, jj, tp, ta, tv );
                    R.torusPos.insert( R.torusPos.end(), tp, tp + 4 );
                    R.torusAxis.insert( R.torusAxis.end(), ta, ta + 4 );
                    R.torusVs.insert( R.torusVs.end(), tv, tv + 4 );
                    R.torusIds.push_back( i );
                    R.torusIds.push_back( j );
                }
            }
        }
        singularities( R, p, track, delta );
        return Rp;
    }

    const char * StatNames = "candidates cover_tests arc_tests arc_updates probe_pairs arc_peak nb_overflow arc_overflow";
} // namespace cbo

extern "C"
{
    // Contour-Buildup oracle build. aabb: min xyz, max xyz (the reference's getAabb); dim: cells per
    // axis (power of two, reference 16). precision 32 or 64.
    void * oracle_cb_build( const float * xyzr, uint32_t n, float probeRadius, const float * aabb, uint32_t dim, uint32_t texSize,
                            int precision, int track, double delta, double dtheta, int threads )
    {
        if ( dim == 0 || ( dim & ( dim - 1 ) ) || texSize == 0 )
            return nullptr;
        cbo::Params p {};
        for ( int a = 0; a < 3; a++ )
        {
            p.origin[ a ] = aabb[ a ];
            p.cell[ a ]   = ( aabb[ 3 + a ] - aabb[ a ] ) / float( dim );
        }
        p.dim = dim, p.rp = probeRadius, p.texSize = texSize;
        if ( precision == 64 )
            return cbo::build<double>( xyzr, n, p, track != 0, delta, dtheta, threads );
        return cbo::build<float>( xyzr, n, p, false, delta, dtheta, threads );
    }

    // Singularity texture of a given probe array (float4 xyz, rp) on the build's grid, FP32:
    // fills texCoord [3 P] (probe order), singTex [3 * 32 P] (sorted probe order) and probeSorted [P].
    int oracle_cb_singularities( void * h, const float * probes, uint32_t P, float * texCoord, float * singTex, uint32_t * probeSorted )
    {
        const cbo::Result * R = static_cast<cbo::Result *>( h );
        cbo::Result         S;
        S.stats.assign( cbo::StatCount, 0 );
        S.probePos.assign( probes, probes + size_t( P ) * 4 );
        cbo::singularities( S, R->prm, false, 0.0 );
        std::memcpy( texCoord, S.texCoord.data(), S.texCoord.size() * 4 );
        std::memcpy( singTex, S.singTex.data(), S.singTex.size() * 4 );
        std::memcpy( probeSorted, S.probeSorted.data(), S.probeSorted.size() * 4 );
        return 0;
    }

    void oracle_cb_free( void * h ) { delete static_cast<cbo::Result *>( h ); }

    const char * oracle_cb_stat_names() { return cbo::StatNames; }

    // Copies array `name` into out (when non-null); returns its element count, sets *elemSize.
    uint64_t oracle_cb_get( void * h, const char * name, void * out, uint32_t * elemSize )
    {
        cbo::Result * R = static_cast<cbo::Result *>( h );
        const std::string nm( name );
        auto put = [ & ]( const auto & v ) -> uint64_t {
            using E = typename std::decay_t<decltype( v )>::value_type;
            if ( elemSize )
                *elemSize = sizeof( E );
            if ( out && !v.empty() )
                std::memcpy( out, v.data(), v.size() * sizeof( E ) );
            return v.size();
        };
        if ( nm == "atoms" ) return put( R->atoms );
        if ( nm == "sorted_to_initial" ) return put( R->sortedToInitial );
        if ( nm == "nb_count0" ) return put( R->nbCount0 );
        if ( nm == "nb_count" ) return put( R->nbCount );
        if ( nm == "nb_ids" ) return put( R->nbIds );
        if ( nm == "removed" ) return put( R->removed );
        if ( nm == "arc_count" ) return put( R->arcCount );
        if ( nm == "visible" ) return put( R->visible );
        if ( nm == "probe_pos" ) return put( R->probePos );
        if ( nm == "probe_pos64" ) return put( R->probePos64 );
        if ( nm == "probe_ids" ) return put( R->probeIds );
        if ( nm == "tria1" ) return put( R->tria1 );
        if ( nm == "tria2" ) return put( R->tria2 );
        if ( nm == "tria3" ) return put( R->tria3 );
        if ( nm == "torus_pos" ) return put( R->torusPos );
        if ( nm == "torus_axis" ) return put( R->torusAxis );
        if ( nm == "torus_vs" ) return put( R->torusVs );
        if ( nm == "torus_ids" ) return put( R->torusIds );
        if ( nm == "probe_sorted" ) return put( R->probeSorted );
        if ( nm == "tex_coord" ) return put( R->texCoord );
        if ( nm == "sing_tex" ) return put( R->singTex );
        if ( nm == "amb_atom" ) return put( R->ambAtom );
        if ( nm == "amb_cover" ) return put( R->ambCover );
        if ( nm == "amb_circle" ) return put( R->ambCircle );
        if ( nm == "amb_probe" ) return put( R->ambProbe );
        if ( nm == "stats" ) return put( R->stats );
        if ( elemSize )
            *elemSize = 0;
        return 0;
    }
}
