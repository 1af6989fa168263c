// Surface components of a complete-surface build (used by the exterior SDF variant).
//
// The boundary of the SAS union A is a closed surface of faces (exposed regions of a sphere), arcs
// (full circles and segments) and vertices (probe positions); each connected piece of it bounds one
// region of the complement of A. Components are formed over nodes
//   [ full circles | segments | vertices | atoms ]  (capacities capFull, capSeg, capRec, n):
// a segment joins its two end vertices, and the loops bounding one face of an atom are joined (an
// atom's exposed sphere may hold several faces, of different components). An atom without arcs is
// its own component when its whole sphere is exposed. Components that bound one region of the
// complement are then joined (see Components); the exterior label is that of the unbounded region.
//
// Point location on a sphere whose faces lie in several components ("mixed" atoms): the geodesic
// from an exposed point u towards the pole of one of the atom's caps (a buried point) leaves the
// face holding u through an arc of that face's boundary, so the first arc it crosses names the
// component.
#pragma once

#include "kernels.h"

namespace sesdf
{
    constexpr uint32_t LabelMixed  = 0xfffffffeu; // atom whose boundary spans components
    constexpr uint32_t LabelNone   = 0xffffffffu; // atom with no exposed boundary (buried)
    constexpr uint32_t MaxAtomArcs = 64;          // arcs per atom grouped into faces in registers
    constexpr uint32_t MaxAtomLoops = 16;

    // Regions of the complement: several components may bound one region (the outer surfaces of
    // separate pieces bound the unbounded one; a piece floating in a cavity shares the cavity's
    // region). A ray from an exposed +x pole runs in the region of the pole's face until it enters
    // the boundary of that region again, or escapes: the two components are joined, or the pole's
    // component joins the node of the unbounded region. The labels are then those of regions.
    // Points farther than rp + band from A are classified by a ray of their own, which only points
    // inside the bounding box of some other region large enough to hold them need (cavMap).
    struct Components
    {
        uint32_t * label;     // per node: component label (compressed union-find parent)
        uint32_t * atomLabel; // per sorted atom: its boundary's label, LabelMixed or LabelNone
        uint8_t *  extStatus; // per sorted atom: 0 no exterior boundary, 1 all of it exterior, 2 part of it
        uint32_t * incCount;  // n: arcs per atom (cursor of the scatter)
        uint32_t * incStart;  // n + 1: exclusive scan of incCount
        uint32_t * incArcs;   // arcs incident to each atom, ascending (arc id = node id)
        uint32_t * extComp;   // [0] exterior label, [1] atoms over the register capacity
        uint32_t * boxes;     // per node: 6 order-preserving float keys (min xyz, max xyz) of its arcs' circles
        uint32_t * cavMap;    // per grid cell: 64 x the largest half-extent (smallest axis) of the boxes of
                              // non-exterior regions covering the cell (0: none)
        uint32_t   capFull, capSeg, capRec, n, cellNb;
    };
    __device__ __forceinline__ uint32_t ext_node( const Components & c ) { return c.capFull + c.capSeg + c.capRec + c.n; }

    __device__ __forceinline__ uint32_t vertex_node( const Components & c, uint32_t v ) { return c.capFull + c.capSeg + v; }
    __device__ __forceinline__ uint32_t atom_node( const Components & c, uint32_t a ) { return c.capFull + c.capSeg + c.capRec + a; }

    // circle (i, j) in world coordinates: centre, unit normal (i -> j), radius
    __device__ __forceinline__ void circle_world( const Buffers & b, uint32_t i, uint32_t j, f3 & ctr, f3 & nrm, float & rad )
    {
        const float4 ai = b.atoms[ i ], aj = b.atoms[ j ];
        const float  Ri = ai.w + b.probeRadius, Rj = aj.w + b.probeRadius;
        const f3     dj = xyz( aj ) - xyz( ai );
        const float  d2 = len2( dj );
        const float  t  = ( Ri * Ri - Rj * Rj + d2 ) / ( 2.f * d2 );
        const f3     cl = dj * t;
        rad             = sqrtf( fmaxf( Ri * Ri - len2( cl ), 0.f ) );
        nrm             = dj * ( 1.0f / sqrtf( d2 ) );
        ctr             = xyz( ai ) + cl;
    }

    __device__ __forceinline__ float angle_ccw( f3 A, f3 u, f3 n )
    {
        float t = atan2f( dot( cross( A, u ), n ), dot( A, u ) );
        return t < 0.f ? t + TwoPi : t;
    }

    __device__ __forceinline__ uint2 arc_atoms( const Buffers & b, const Components & c, uint32_t e )
    {
        if ( e < c.capFull )
            return b.fullCircles[ e ];
        const uint4 s = b.segments[ e - c.capFull ];
        return make_uint2( s.x, s.y );
    }

    // cap of circle (a, o) on sphere a: pole (unit, towards c_o) and the cosine of its half angle
    __device__ __forceinline__ void cap_of( const Buffers & b, uint32_t a, uint32_t o, f3 & pole, float & cosT )
    {
        const float4 A = b.atoms[ a ], O = b.atoms[ o ];
        const float  Ra = A.w + b.probeRadius, Ro = O.w + b.probeRadius;
        const f3     d  = xyz( O ) - xyz( A );
        const float  dd = sqrtf( len2( d ) );
        pole            = d * ( 1.f / dd );
        cosT            = ( Ra * Ra - Ro * Ro + dd * dd ) / ( 2.f * dd * Ra );
    }

    // x (world, on the circle of arc e) lies on the arc: segments run from A counter-clockwise about
    // n(i -> j) to B
    __device__ __forceinline__ bool on_arc( const Buffers & b, const Components & c, uint32_t e, f3 x )
    {
        if ( e < c.capFull )
            return true;
        const uint4 s = b.segments[ e - c.capFull ];
        f3          ctr, nrm;
        float       rad;
        circle_world( b, s.x, s.y, ctr, nrm, rad );
        const f3 A = xyz( b.concavePos[ s.z ] ) - ctr, B = xyz( b.concavePos[ s.w ] ) - ctr;
        return angle_ccw( A, x - ctr, nrm ) <= angle_ccw( A, B, nrm );
    }

    // geodesic frame on the unit sphere from v1 towards v2: v(phi) = cos phi v1 + sin phi w, phi in
    // [0, Phi]; false when v1 and v2 are (nearly) antipodal
    __device__ __forceinline__ bool geo_frame( f3 v1, f3 v2, f3 & w, float & Phi )
    {
        const float cd = dot( v1, v2 );
        const float s  = sqrtf( len2( cross( v1, v2 ) ) );
        if ( s < 1e-5f && cd < 0.f )
            return false;
        const f3    t  = v2 - v1 * cd;
        const float tl = sqrtf( len2( t ) );
        w              = tl > 0.f ? t * ( 1.f / tl ) : mk( 0.f, 0.f, 0.f );
        Phi            = atan2f( s, cd );
        return true;
    }

    // Angles in [0, Phi] where the geodesic (v1, w, Phi) on sphere a (centre ca, radius Ra) meets
    // arc e; returns how many (0..2), angles in phis.
    __device__ __forceinline__ int cross_arc( const Buffers & b, const Components & c, uint32_t a, f3 ca, float Ra, f3 v1, f3 w, float Phi,
                                              uint32_t e, float * phis )
    {
        const uint2 uv = arc_atoms( b, c, e );
        f3          pole;
        float       cosT;
        cap_of( b, a, uv.x == a ? uv.y : uv.x, pole, cosT );
        const float p = dot( v1, pole ), q = dot( w, pole ), rho = sqrtf( p * p + q * q );
        if ( !( rho > fabsf( cosT ) ) )
            return 0;
        const float ph0 = atan2f( q, p ), dph = acosf( fminf( fmaxf( cosT / rho, -1.f ), 1.f ) );
        int         cnt = 0;
        for ( int k = 0; k < 2; k++ )
        {
            float ph = fmodf( ( k ? ph0 + dph : ph0 - dph ) + 2.f * TwoPi, TwoPi );
            if ( ph > Phi )
                continue;
            const f3 v = v1 * cosf( ph ) + w * sinf( ph );
            if ( on_arc( b, c, e, ca + v * Ra ) )
                phis[ cnt++ ] = ph;
        }
        return cnt;
    }

    // First arc of atom a crossed by the geodesic from u towards `target` (unit vectors from c_a);
    // LabelNone if none is crossed.
    __device__ __forceinline__ uint32_t first_hit_towards( const Buffers & b, const Components & c, uint32_t a, f3 u, f3 target )
    {
        const float4   A  = b.atoms[ a ];
        const f3       ca = xyz( A );
        const float    Ra = A.w + b.probeRadius;
        f3             w;
        float          Phi;
        if ( !geo_frame( u, target, w, Phi ) )
            return LabelNone;
        float    best = 3.0e38f;
        uint32_t arc  = LabelNone;
        for ( uint32_t t2 = c.incStart[ a ]; t2 < c.incStart[ a + 1 ]; t2++ )
        {
            float          ph[ 2 ];
            const uint32_t e = c.incArcs[ t2 ];
            const int      k = cross_arc( b, c, a, ca, Ra, u, w, Phi, e, ph );
            for ( int q = 0; q < k; q++ )
                if ( ph[ q ] < best )
                    best = ph[ q ], arc = e;
        }
        return arc;
    }

    // First arc of atom a crossed from the exposed point u (unit vector from c_a) towards a cap
    // pole; LabelNone if none is found (u not exposed).
    __device__ __forceinline__ uint32_t first_hit( const Buffers & b, const Components & c, uint32_t a, f3 u )
    {
        const float4   A  = b.atoms[ a ];
        const f3       ca = xyz( A );
        const float    Ra = A.w + b.probeRadius;
        const uint32_t t0 = c.incStart[ a ], t1 = c.incStart[ a + 1 ];
        for ( uint32_t t = t0; t < t1; t++ )
        {
            const uint2 uv = arc_atoms( b, c, c.incArcs[ t ] );
            f3          pole, w;
            float       cosT, Phi;
            cap_of( b, a, uv.x == a ? uv.y : uv.x, pole, cosT );
            if ( !geo_frame( u, pole, w, Phi ) )
                continue;
            float    best = 3.0e38f;
            uint32_t arc  = LabelNone;
            for ( uint32_t t2 = t0; t2 < t1; t2++ )
            {
                float          ph[ 2 ];
                const uint32_t e = c.incArcs[ t2 ];
                const int      k = cross_arc( b, c, a, ca, Ra, u, w, Phi, e, ph );
                for ( int q = 0; q < k; q++ )
                    if ( ph[ q ] < best )
                        best = ph[ q ], arc = e;
            }
            return arc;
        }
        return LabelNone;
    }

    // First SAS sphere (other than `skip`) entered by the ray p + t x^, t > 0: its atom (LabelNone if
    // the ray escapes) and the entry x. Spheres that can meet the ray have centres in the rows of
    // cells within R_max of it, in the columns from p's on (single-molecule grids).
    __device__ __forceinline__ uint32_t ray_hit( const Buffers & b, f3 p, uint32_t skip, float & hitX )
    {
        const Molecule M    = b.mols[ 0 ];
        const float    Rmax = __uint_as_float( b.ctl->radiusMaxBits ) + b.probeRadius;
        int            gx, gy, gz;
        cell_of( M, p.x, p.y, p.z, gx, gy, gz );
        const int r    = max( int( ceilf( ( Rmax + 1e-3f ) / M.cell ) ), 1 );
        uint32_t  best = LabelNone;
        hitX           = 3.0e38f;
        for ( int z = max( gz - r, 0 ); z <= min( gz + r, M.dz - 1 ); z++ )
            for ( int y = max( gy - r, 0 ); y <= min( gy + r, M.dy - 1 ); y++ )
            {
                const uint32_t a0 = b.cellStart[ cell_index( M, gx, y, z ) ], a1 = b.cellStart[ cell_index( M, M.dx - 1, y, z ) + 1 ];
                for ( uint32_t a = a0; a < a1; a++ )
                {
                    const float4 A  = b.atoms[ a ];
                    const float  R  = A.w + b.probeRadius;
                    const float  dy = A.y - p.y, dz = A.z - p.z;
                    const float  h2 = R * R - dy * dy - dz * dz;
                    if ( a == skip || !( h2 > 0.f ) )
                        continue;
                    const float x = A.x - sqrtf( h2 );
                    if ( x > p.x && x < hitX )
                        hitX = x, best = a;
                }
            }
        return best;
    }

    // Component of the exposed point c_a + R_a u of sphere a.
    __device__ __forceinline__ uint32_t label_at( const Buffers & b, const Components & c, uint32_t a, f3 u )
    {
        const uint32_t l = c.atomLabel[ a ];
        if ( l != LabelMixed )
            return l;
        const uint32_t e = first_hit( b, c, a, u );
        return e == LabelNone ? LabelNone : c.label[ e ];
    }

    // Points outside A and farther than rp + band from it (exterior variant): true when the point
    // lies in a region other than the unbounded one (a cavity), by its ray.
    __device__ __forceinline__ bool far_point_in_cavity( const Buffers & b, const Components & c, f3 q, float D, uint32_t extC )
    {
        int gx, gy, gz;
        cell_of( b.mols[ 0 ], q.x, q.y, q.z, gx, gy, gz );
        if ( float( c.cavMap[ cell_index( b.mols[ 0 ], gx, gy, gz ) ] ) < D * 64.f )
            return false; // no region but the unbounded one can hold a ball of radius D around q
        float          hx;
        const uint32_t h = ray_hit( b, q, LabelNone, hx );
        if ( h == LabelNone )
            return false;
        const float4   A  = b.atoms[ h ];
        const uint32_t lh = label_at( b, c, h, unit( mk( hx, q.y, q.z ) - xyz( A ) ) );
        return lh != LabelNone && lh != extC;
    }
} // namespace sesdf
