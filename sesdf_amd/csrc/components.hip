// Surface components of a complete-surface build: the kernels that label them (exterior SDF).
//
// No reference counterpart: the reference has no SDF and no exterior classification of cavities;
// its exterior path (bcs::Ssesdf) builds a different element set. The rule is set out in
// components.h. Once per build, on the handle's stream:
//   init      labels = identity, per-atom arc counts zeroed, argmax of x + r;
//   count     arcs per atom; every segment joined to its two end vertices (lock-free union-find);
//   scan      incident-arc offsets per atom (tile scan, self-clearing counts);
//   scatter   incident arcs per atom;
//   faces     per atom: loops (arcs chained at shared vertices) and, with two or more loops, which
//             of them bound one face (joined);
//   compress, atom labels;
//   poles     a ray from every exposed +x pole joins the pole's component with the one it enters
//             next, or with the unbounded region's node (regions, see components.h);
//   compress, atom labels, exterior label, per-atom exterior status;
//   boxes     per region: box of its arcs' circles; cells covered by the box of a non-exterior
//             region, with the box's smallest half-extent (the far-point ray filter).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "components.h"
#include "handle.h"

namespace sesdf
{
    // parent[] is read with agent-scope atomic loads: other XCDs hook through atomics and a plain
    // load could be served from this XCD's L2 with a stale value.
    __device__ __forceinline__ uint32_t cf_load( uint32_t * parent, uint32_t a )
    {
        return __hip_atomic_load( parent + a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT );
    }
    __device__ __forceinline__ uint32_t cf_find( uint32_t * parent, uint32_t a )
    {
        uint32_t p = cf_load( parent, a );
        while ( p != a )
        {
            a = p;
            p = cf_load( parent, a );
        }
        return a;
    }
    // Union by hooking the larger root under the smaller with a compare-and-swap on the root; a
    // failed swap means the root was hooked meanwhile, and the walk starts again.
    __device__ __forceinline__ void cf_union( uint32_t * parent, uint32_t u, uint32_t v )
    {
        for ( ;; )
        {
            u = cf_find( parent, u );
            v = cf_find( parent, v );
            if ( u == v )
                return;
            const uint32_t hi = max( u, v ), lo = min( u, v );
            if ( atomicCAS( parent + hi, hi, lo ) == hi )
                return;
        }
    }

    // floats as order-preserving unsigned keys
    __device__ __forceinline__ uint32_t fkey( float v )
    {
        const uint32_t bits = __float_as_uint( v );
        return ( bits & 0x80000000u ) ? ~bits : ( bits | 0x80000000u );
    }
    __device__ __forceinline__ float fkey_inv( uint32_t k ) { return __uint_as_float( ( k & 0x80000000u ) ? ( k & 0x7fffffffu ) : ~k ); }

    // labels = identity over the nodes (and the unbounded region's node), empty boxes, zero counts
    // and cell map
    __global__ void comp_init_kernel( Components c, uint32_t nodes )
    {
        const uint32_t stride = gridDim.x * blockDim.x;
        for ( uint32_t x = blockIdx.x * blockDim.x + threadIdx.x; x < nodes; x += stride )
        {
            c.label[ x ] = x;
            for ( int k = 0; k < 3; k++ )
                c.boxes[ 6 * size_t( x ) + k ] = 0xffffffffu, c.boxes[ 6 * size_t( x ) + 3 + k ] = 0u;
        }
        for ( uint32_t a = blockIdx.x * blockDim.x + threadIdx.x; a < c.n; a += stride )
            c.incCount[ a ] = 0;
        for ( uint32_t q = blockIdx.x * blockDim.x + threadIdx.x; q < c.cellNb; q += stride )
            c.cavMap[ q ] = 0;
        if ( blockIdx.x == 0 && threadIdx.x == 0 )
            c.extComp[ 0 ] = 0, c.extComp[ 1 ] = 0;
    }

    // arcs in build order: full circles [0, Cf), then segments (arc id capFull + s)
    __global__ void comp_count_kernel( Buffers b, Components c )
    {
        const uint32_t Cf = min( b.ctl->fullCircleNb, c.capFull ), G = min( b.ctl->segmentNb, c.capSeg );
        for ( uint32_t q = blockIdx.x * blockDim.x + threadIdx.x; q < Cf + G; q += gridDim.x * blockDim.x )
        {
            const uint32_t e  = q < Cf ? q : c.capFull + ( q - Cf );
            const uint2    uv = arc_atoms( b, c, e );
            atomicAdd( c.incCount + uv.x, 1u );
            atomicAdd( c.incCount + uv.y, 1u );
            if ( q >= Cf )
            {
                const uint4 s = b.segments[ q - Cf ];
                cf_union( c.label, e, vertex_node( c, s.z ) );
                cf_union( c.label, e, vertex_node( c, s.w ) );
            }
        }
    }

    __global__ void comp_scatter_kernel( Buffers b, Components c )
    {
        const uint32_t Cf = min( b.ctl->fullCircleNb, c.capFull ), G = min( b.ctl->segmentNb, c.capSeg );
        for ( uint32_t q = blockIdx.x * blockDim.x + threadIdx.x; q < Cf + G; q += gridDim.x * blockDim.x )
        {
            const uint32_t e  = q < Cf ? q : c.capFull + ( q - Cf );
            const uint2    uv = arc_atoms( b, c, e );
            c.incArcs[ c.incStart[ uv.x ] + atomicAdd( c.incCount + uv.x, 1u ) ] = e;
            c.incArcs[ c.incStart[ uv.y ] + atomicAdd( c.incCount + uv.y, 1u ) ] = e;
        }
    }

    // A point of arc e just off it on the exposed side of sphere a (unit vector from c_a): the arc's
    // midpoint moved by `delta` radians away from the cap the arc bounds. The point lies in the face
    // the arc bounds only if no other arc passes between it and the arc (a small loop, or another
    // arc close by): then the geodesic from it back towards the cap's pole crosses e first. The
    // offset is shrunk until that holds (1e-3, 1e-4, 1e-5 rad); `ok` tells whether it did.
    __device__ f3 inside_point_checked( const Buffers & b, const Components & c, uint32_t a, uint32_t e, bool & ok );
    __device__ f3 inside_point( const Buffers & b, const Components & c, uint32_t a, uint32_t e, float delta )
    {
        const uint2 uv = arc_atoms( b, c, e );
        f3          ctr, nrm;
        float       rad;
        circle_world( b, uv.x, uv.y, ctr, nrm, rad );
        f3 x;
        if ( e < c.capFull )
        {
            const f3 ref = fabsf( nrm.x ) < 0.9f ? mk( 1.f, 0.f, 0.f ) : mk( 0.f, 1.f, 0.f );
            x            = ctr + unit( cross( nrm, ref ) ) * rad;
        }
        else
        {
            const uint4 s   = b.segments[ e - c.capFull ];
            const f3    A   = xyz( b.concavePos[ s.z ] ) - ctr, B = xyz( b.concavePos[ s.w ] ) - ctr;
            float       ang = atan2f( dot( cross( A, B ), nrm ), dot( A, B ) );
            if ( ang <= 0.f )
                ang += TwoPi;
            const float h = 0.5f * ang;
            x             = ctr + A * cosf( h ) + cross( nrm, A ) * sinf( h );
        }
        f3    pole;
        float cosT;
        cap_of( b, a, uv.x == a ? uv.y : uv.x, pole, cosT );
        const f3 v = unit( x - xyz( b.atoms[ a ] ) );
        const f3 t = unit( v * dot( v, pole ) - pole ); // tangent, away from the cap
        return unit( v + t * delta );
    }
    __device__ f3 inside_point_checked( const Buffers & b, const Components & c, uint32_t a, uint32_t e, bool & ok )
    {
        const uint2 uv = arc_atoms( b, c, e );
        f3          pole;
        float       cosT;
        cap_of( b, a, uv.x == a ? uv.y : uv.x, pole, cosT );
        f3 x = mk( 0.f, 0.f, 0.f );
        ok   = false;
        for ( float delta = 1e-3f; delta > 5e-6f && !ok; delta *= 0.1f )
        {
            x  = inside_point( b, c, a, e, delta );
            ok = first_hit_towards( b, c, a, x, pole ) == e;
        }
        return x;
    }

    // Loops and faces of each atom. Arcs of atom a chain at shared vertices into loops; two loops
    // bound the same face iff the geodesic between points just inside them crosses every loop of
    // the atom an even number of times (regions of the sphere cut by disjoint closed curves are told
    // apart by the side of each curve they lie on).
    __global__ void __launch_bounds__( 64 ) comp_faces_kernel( Buffers b, Components c )
    {
        const uint32_t a = blockIdx.x * blockDim.x + threadIdx.x;
        if ( a >= c.n )
            return;
        const uint32_t t0 = c.incStart[ a ], m = c.incStart[ a + 1 ] - t0;
        if ( m < 2 )
            return;
        if ( m > MaxAtomArcs )
        {
            // beyond the register capacity the atom's loops are all joined (counted)
            atomicAdd( c.extComp + 1, 1u );
            for ( uint32_t t = 1; t < m; t++ )
                cf_union( c.label, c.incArcs[ t0 ], c.incArcs[ t0 + t ] );
            return;
        }
        uint32_t arcs[ MaxAtomArcs ];
        uint8_t  lp[ MaxAtomArcs ], root0[ MaxAtomArcs ], par[ MaxAtomArcs ];
        for ( uint32_t x = 0; x < m; x++ )
        {
            // insertion sort: the scatter's order depends on atomics
            const uint32_t e = c.incArcs[ t0 + x ];
            uint32_t       y = x;
            while ( y > 0 && arcs[ y - 1 ] > e )
            {
                arcs[ y ] = arcs[ y - 1 ];
                y--;
            }
            arcs[ y ] = e;
            lp[ x ]   = uint8_t( x );
        }
        for ( uint32_t x = 0; x < m; x++ )
            c.incArcs[ t0 + x ] = arcs[ x ];
        auto lf = [ & ]( uint32_t x ) {
            while ( lp[ x ] != x )
                x = lp[ x ] = lp[ lp[ x ] ];
            return x;
        };
        auto lu = [ & ]( uint32_t x, uint32_t y ) {
            x = lf( x ), y = lf( y );
            if ( x != y )
                lp[ max( x, y ) ] = uint8_t( min( x, y ) );
        };
        for ( uint32_t x = 0; x < m; x++ )
        {
            if ( arcs[ x ] < c.capFull )
                continue;
            const uint4 sx = b.segments[ arcs[ x ] - c.capFull ];
            for ( uint32_t y = x + 1; y < m; y++ )
            {
                if ( arcs[ y ] < c.capFull )
                    continue;
                const uint4 sy = b.segments[ arcs[ y ] - c.capFull ];
                if ( sx.z == sy.z || sx.z == sy.w || sx.w == sy.z || sx.w == sy.w )
                    lu( x, y );
            }
        }
        uint32_t reps[ MaxAtomLoops ];
        uint32_t L = 0;
        for ( uint32_t x = 0; x < m; x++ )
        {
            root0[ x ] = uint8_t( lf( x ) );
            if ( root0[ x ] == x )
            {
                if ( L < MaxAtomLoops )
                    reps[ L ] = x;
                L++;
            }
        }
        if ( L < 2 )
            return;
        if ( L > MaxAtomLoops )
        {
            atomicAdd( c.extComp + 1, 1u );
            for ( uint32_t t = 1; t < m; t++ )
                cf_union( c.label, arcs[ 0 ], arcs[ t ] );
            return;
        }
        const float4 A  = b.atoms[ a ];
        const f3     ca = xyz( A );
        const float  Ra = A.w + b.probeRadius;
        f3           pts[ MaxAtomLoops ];
        for ( uint32_t p = 0; p < L; p++ )
        {
            // any arc of the loop bounds the same face: the first whose inside point checks out
            bool ok = false;
            for ( uint32_t x = 0; x < m && !ok; x++ )
                if ( root0[ x ] == reps[ p ] )
                    pts[ p ] = inside_point_checked( b, c, a, arcs[ x ], ok );
            if ( !ok )
                atomicAdd( c.extComp + 1, 1u ); // counted with the atoms over capacity
        }
        for ( uint32_t p = 0; p < L; p++ )
            for ( uint32_t q = p + 1; q < L; q++ )
            {
                if ( lf( reps[ p ] ) == lf( reps[ q ] ) )
                    continue;
                for ( uint32_t x = 0; x < m; x++ )
                    par[ x ] = 0;
                // geodesic p -> q, through a third point when they are nearly antipodal
                f3    path[ 3 ] = { pts[ p ], pts[ q ], pts[ q ] };
                int   legs      = 1;
                f3    w;
                float Phi;
                if ( !geo_frame( pts[ p ], pts[ q ], w, Phi ) )
                {
                    const f3 ref = fabsf( pts[ p ].x ) < 0.9f ? mk( 1.f, 0.f, 0.f ) : mk( 0.f, 1.f, 0.f );
                    path[ 1 ]    = unit( cross( pts[ p ], ref ) );
                    legs         = 2;
                }
                for ( int l = 0; l < legs; l++ )
                {
                    geo_frame( path[ l ], path[ l + 1 ], w, Phi );
                    for ( uint32_t x = 0; x < m; x++ )
                    {
                        float ph[ 2 ];
                        par[ root0[ x ] ] += uint8_t( cross_arc( b, c, a, ca, Ra, path[ l ], w, Phi, arcs[ x ], ph ) );
                    }
                }
                bool even = true;
                for ( uint32_t r = 0; r < L; r++ )
                    even = even && ( par[ reps[ r ] ] & 1 ) == 0;
                if ( even )
                {
                    lu( reps[ p ], reps[ q ] );
                    cf_union( c.label, arcs[ reps[ p ] ], arcs[ reps[ q ] ] );
                }
            }
    }

    __global__ void comp_compress_kernel( Components c, uint32_t nodes )
    {
        for ( uint32_t x = blockIdx.x * blockDim.x + threadIdx.x; x < nodes; x += gridDim.x * blockDim.x )
            c.label[ x ] = cf_find( c.label, x );
    }

    // Label of each atom's boundary: the label of its arcs when they agree (LabelMixed otherwise).
    // An atom without arcs has a boundary-free exposed region: all of its sphere (no neighbour cuts a
    // cap from it or contains it: its own component) or nothing (buried).
    __global__ void comp_atom_kernel( Buffers b, Components c )
    {
        const uint32_t a = blockIdx.x * blockDim.x + threadIdx.x;
        if ( a >= c.n )
            return;
        const uint32_t t0 = c.incStart[ a ], t1 = c.incStart[ a + 1 ];
        if ( t0 == t1 )
        {
            const float4 A  = b.atoms[ a ];
            const float  Ra = A.w + b.probeRadius;
            bool         buried = false;
            for ( uint32_t l = 0; l < b.nbCount[ a ] && !buried; l++ )
            {
                const float4 J  = b.atoms[ b.nbIds[ size_t( a ) * b.K + l ] ];
                const float  Rj = J.w + b.probeRadius;
                const float  d  = sqrtf( len2( xyz( J ) - xyz( A ) ) );
                buried          = d < Ra + Rj && d + Rj > Ra;
            }
            c.atomLabel[ a ] = buried ? LabelNone : c.label[ atom_node( c, a ) ];
            return;
        }
        uint32_t l = c.label[ c.incArcs[ t0 ] ];
        for ( uint32_t t = t0 + 1; t < t1; t++ )
            if ( c.label[ c.incArcs[ t ] ] != l )
                l = LabelMixed;
        c.atomLabel[ a ] = l;
    }

    // Rays from the exposed +x poles (see Components).
    __global__ void comp_pole_kernel( Buffers b, Components c )
    {
        const uint32_t a = blockIdx.x * blockDim.x + threadIdx.x;
        if ( a >= c.n )
            return;
        const float4 A    = b.atoms[ a ];
        const float  Ra   = A.w + b.probeRadius;
        const f3     pole = mk( A.x + Ra, A.y, A.z );
        for ( uint32_t l = 0; l < b.nbCount[ a ]; l++ )
        {
            const float4 J  = b.atoms[ b.nbIds[ size_t( a ) * b.K + l ] ];
            const float  Rj = J.w + b.probeRadius;
            if ( len2( pole - xyz( J ) ) < Rj * Rj )
                return; // the pole is buried
        }
        const uint32_t l0 = label_at( b, c, a, mk( 1.f, 0.f, 0.f ) );
        if ( l0 == LabelNone )
            return;
        float          hx;
        const uint32_t h  = ray_hit( b, pole, a, hx );
        uint32_t       l1 = ext_node( c );
        if ( h != LabelNone )
            l1 = label_at( b, c, h, unit( mk( hx, pole.y, pole.z ) - xyz( b.atoms[ h ] ) ) );
        if ( l1 != LabelNone )
            cf_union( c.label, l0, l1 );
    }

    __global__ void comp_ext_kernel( Components c )
    {
        if ( threadIdx.x == 0 )
            c.extComp[ 0 ] = c.label[ ext_node( c ) ];
    }

    // Box of every non-exterior region: the boxes of its arcs' circles (the region's extreme points
    // lie on arcs or vertices: a face bulges into the atom, away from the region).
    __global__ void comp_box_kernel( Buffers b, Components c )
    {
        const uint32_t Cf = min( b.ctl->fullCircleNb, c.capFull ), G = min( b.ctl->segmentNb, c.capSeg );
        const uint32_t ext = c.extComp[ 0 ];
        for ( uint32_t q = blockIdx.x * blockDim.x + threadIdx.x; q < Cf + G; q += gridDim.x * blockDim.x )
        {
            const uint32_t e = q < Cf ? q : c.capFull + ( q - Cf );
            const uint32_t l = c.label[ e ];
            if ( l == ext )
                continue;
            const uint2 uv = arc_atoms( b, c, e );
            f3          ctr, nrm;
            float       rad;
            circle_world( b, uv.x, uv.y, ctr, nrm, rad );
            const float ex[ 3 ] = { rad * sqrtf( fmaxf( 1.f - nrm.x * nrm.x, 0.f ) ), rad * sqrtf( fmaxf( 1.f - nrm.y * nrm.y, 0.f ) ),
                                    rad * sqrtf( fmaxf( 1.f - nrm.z * nrm.z, 0.f ) ) };
            const float ce[ 3 ] = { ctr.x, ctr.y, ctr.z };
            for ( int k = 0; k < 3; k++ )
            {
                // 1e-3 A of slack covers the rounding of the circle's extent
                atomicMin( c.boxes + 6 * size_t( l ) + k, fkey( ce[ k ] - ex[ k ] - 1e-3f ) );
                atomicMax( c.boxes + 6 * size_t( l ) + 3 + k, fkey( ce[ k ] + ex[ k ] + 1e-3f ) );
            }
        }
    }

    // A point farther than D > rp from A can lie in a bounded region only if the region's box holds
    // a ball of radius D around it: cells covered by such boxes get the box's smallest half-extent.
    __global__ void comp_cavmap_kernel( Buffers b, Components c, uint32_t nodes )
    {
        const Molecule M = b.mols[ 0 ];
        for ( uint32_t l = blockIdx.x * blockDim.x + threadIdx.x; l < nodes; l += gridDim.x * blockDim.x )
        {
            if ( c.label[ l ] != l || l == c.extComp[ 0 ] || c.boxes[ 6 * size_t( l ) ] == 0xffffffffu )
                continue;
            float lo[ 3 ], hi[ 3 ], half = 3.0e38f;
            for ( int k = 0; k < 3; k++ )
            {
                lo[ k ] = fkey_inv( c.boxes[ 6 * size_t( l ) + k ] ), hi[ k ] = fkey_inv( c.boxes[ 6 * size_t( l ) + 3 + k ] );
                half    = fminf( half, 0.5f * ( hi[ k ] - lo[ k ] ) );
            }
            if ( !( half > b.probeRadius ) )
                continue;
            const uint32_t v = uint32_t( fminf( ceilf( half * 64.f ), 4.0e9f ) );
            int            x0, y0, z0, x1, y1, z1;
            cell_of( M, lo[ 0 ], lo[ 1 ], lo[ 2 ], x0, y0, z0 );
            cell_of( M, hi[ 0 ], hi[ 1 ], hi[ 2 ], x1, y1, z1 );
            for ( int z = z0; z <= z1; z++ )
                for ( int y = y0; y <= y1; y++ )
                    for ( int x = x0; x <= x1; x++ )
                        atomicMax( c.cavMap + cell_index( M, x, y, z ), v );
        }
    }

    __global__ void comp_status_kernel( Components c )
    {
        const uint32_t a = blockIdx.x * blockDim.x + threadIdx.x;
        if ( a >= c.n )
            return;
        const uint32_t l = c.atomLabel[ a ], ext = c.extComp[ 0 ];
        uint8_t        st = l == ext ? 1 : 0;
        if ( l == LabelMixed )
        {
            bool any = false, all = true;
            for ( uint32_t t = c.incStart[ a ]; t < c.incStart[ a + 1 ]; t++ )
            {
                const bool x = c.label[ c.incArcs[ t ] ] == ext;
                any = any || x, all = all && x;
            }
            st = all ? 1 : any ? 2 : 0;
        }
        c.extStatus[ a ] = st;
    }

    sesdf_status prepare_components( Handle & h, bool mayAlloc )
    {
        SdfState & d = h.sdf;
        if ( d.compAlloc && d.compSerial == h.buildSerial )
            return SESDF_OK;
        const uint32_t n = h.n;
        if ( !d.compAlloc || d.comp.capFull < h.cap.full || d.comp.capSeg < h.cap.segments || d.comp.capRec < h.cap.records ||
             d.comp.cellNb != h.cellNb )
        {
            if ( !mayAlloc )
                return SESDF_ERR_CAPACITY;
            if ( d.compAlloc )
            {
                h.sdfAllocs.erase( std::find( h.sdfAllocs.begin(), h.sdfAllocs.end(), d.compAlloc ) );
                (void)hipFree( d.compAlloc );
                d.compAlloc = nullptr;
            }
            const uint64_t nodes = uint64_t( h.cap.full ) + h.cap.segments + h.cap.records + n + 1;
            const uint64_t arcs  = uint64_t( h.cap.full ) + h.cap.segments;
            if ( nodes >= 0xfffffff0ull )
                return SESDF_ERR_CAPACITY;
            auto         up    = []( size_t v ) { return ( v + 255 ) & ~size_t( 255 ); };
            const size_t bytes = 256 + up( nodes * 4 ) + up( nodes * 24 ) + up( size_t( h.cellNb ) * 4 ) + up( size_t( n ) * 4 ) +
                                 2 * up( ( size_t( n ) + 1 ) * 4 ) + up( 2 * arcs * 4 ) + up( n + 1 );
            void *       p     = nullptr;
            if ( hipMalloc( &p, bytes ) != hipSuccess )
                return SESDF_ERR_HIP;
            h.sdfAllocs.push_back( p );
            h.sdfBytes += bytes;
            d.compAlloc       = p;
            char * q          = static_cast<char *>( p );
            d.comp.extComp    = reinterpret_cast<uint32_t *>( q ), q += 256;
            d.comp.label      = reinterpret_cast<uint32_t *>( q ), q += up( nodes * 4 );
            d.comp.boxes      = reinterpret_cast<uint32_t *>( q ), q += up( nodes * 24 );
            d.comp.cavMap     = reinterpret_cast<uint32_t *>( q ), q += up( size_t( h.cellNb ) * 4 );
            d.comp.atomLabel  = reinterpret_cast<uint32_t *>( q ), q += up( size_t( n ) * 4 );
            d.comp.incCount   = reinterpret_cast<uint32_t *>( q ), q += up( ( size_t( n ) + 1 ) * 4 );
            d.comp.incStart   = reinterpret_cast<uint32_t *>( q ), q += up( ( size_t( n ) + 1 ) * 4 );
            d.comp.incArcs    = reinterpret_cast<uint32_t *>( q ), q += up( 2 * arcs * 4 );
            d.comp.extStatus  = reinterpret_cast<uint8_t *>( q );
            d.comp.capFull    = h.cap.full;
            d.comp.capSeg     = h.cap.segments;
            d.comp.capRec     = h.cap.records;
            d.comp.cellNb     = h.cellNb;
        }
        d.comp.n              = n;
        const Components c    = d.comp;
        const uint32_t   nodes = c.capFull + c.capSeg + c.capRec + n + 1;
        hipStream_t      s    = h.stream;
        const uint32_t   wide = uint32_t( device_cu_count() ) * 8;
        {
            const uint32_t items = std::max( nodes, h.cellNb );
            comp_init_kernel<<<std::min<uint32_t>( ( items + 255 ) / 256, wide ), 256, 0, s>>>( c, nodes );
        }
        if ( n )
        {
            const uint32_t arcBlocks = std::min<uint32_t>( ( c.capFull + c.capSeg + 255 ) / 256, wide );
            comp_count_kernel<<<std::max( arcBlocks, 1u ), 256, 0, s>>>( h.b, c );
            ScanArgs a {};
            a.in[ 0 ]     = c.incCount;
            a.out[ 0 ]    = c.incStart;
            a.n           = n;
            a.ticket      = TicketComp;
            a.stage       = 7;
            a.clearInput  = 1; // the counts become the scatter's cursors
            launch_scan<1>( h.b, a, s );
            comp_scatter_kernel<<<std::max( arcBlocks, 1u ), 256, 0, s>>>( h.b, c );
            comp_faces_kernel<<<( n + 63 ) / 64, 64, 0, s>>>( h.b, c );
            const uint32_t nodeBlocks = std::min<uint32_t>( ( nodes + 255 ) / 256, wide );
            comp_compress_kernel<<<nodeBlocks, 256, 0, s>>>( c, nodes );
            comp_atom_kernel<<<( n + 255 ) / 256, 256, 0, s>>>( h.b, c );
            comp_pole_kernel<<<( n + 255 ) / 256, 256, 0, s>>>( h.b, c );
            comp_compress_kernel<<<nodeBlocks, 256, 0, s>>>( c, nodes );
            comp_atom_kernel<<<( n + 255 ) / 256, 256, 0, s>>>( h.b, c );
            comp_ext_kernel<<<1, 64, 0, s>>>( c );
            comp_status_kernel<<<( n + 255 ) / 256, 256, 0, s>>>( c );
            comp_box_kernel<<<std::max( arcBlocks, 1u ), 256, 0, s>>>( h.b, c );
            comp_cavmap_kernel<<<nodeBlocks, 256, 0, s>>>( h.b, c, nodes );
        }
        if ( hipGetLastError() != hipSuccess )
            return SESDF_ERR_HIP;
        d.compSerial = h.buildSerial;
        return SESDF_OK;
    }
} // namespace sesdf
