// SDF sampling over the built complete surface, and its gradient with the topology held fixed.
//
// The reference ships no SDF evaluator (its renderer is not in the repository); this library
// defines  SDF(q) = clamp(rp + f(q), -band, band)  with f the signed distance to the boundary of
// the SAS union A (negative inside). Boundary elements are those of the complete-surface build:
//   faces    : sphere a where the radial projection of q is not inside another sphere of N(a);
//   arcs     : full circles and segments (A -> B counter-clockwise about n(i -> j));
//   vertices : concave patch (probe) positions.
// Outside A, f = min_a |q - c_a| - R_a. The field is 1-Lipschitz and its zero set is the SES.
// Exterior variant: only elements of the exterior surface component count, and points outside A
// whose nearest sphere lies in another component (probe-accessible cavities) are inside.
//
// Gradient (topology fixed): with p the nearest boundary point, e = (p - q) / |p - q| and nu
// solving sum_m nu_m (p - c_m) = e over the sphere constraints |p - c_m| = R_m of the winning
// element (1 for a face, 2 for an arc, 3 for a vertex):
//   dSDF/dc_m = -nu_m (p - c_m),  dSDF/dr_m = -nu_m R_m   (envelope theorem);
// outside A: dSDF/dc_a = -(q - c_a) / |q - c_a|, dSDF/dr_a = -1.
#include <algorithm>
#include <cstring>
#include <vector>

#include "handle.h"

namespace sesdf
{
    // winner code: type << 40 | index (index: sorted atom, full circle, segment or concave record)
    enum SdfType : uint32_t
    {
        SdfOutside = 0,
        SdfFace    = 1,
        SdfCircle  = 2,
        SdfSegment = 3,
        SdfVertex  = 4,
        SdfClamped = 5
    };
    __device__ __forceinline__ uint64_t code_of( uint32_t type, uint32_t idx ) { return ( uint64_t( type ) << 40 ) | idx; }

    struct SdfArgs
    {
        const float4 *   points;     // n points (xyz_) or null: grid
        float            ox, oy, oz, sx, sy, sz;
        uint32_t         dx, dy, dz;
        uint64_t         n;
        float            band;
        int              exterior;
        const uint32_t * comp;       // surface component per sorted atom (exterior)
        const uint32_t * extComp;    // label of the exterior component (device)
        float *          out;
        uint64_t *       codes;      // optional winner codes
        const float *    weights;    // backward
        int              recompute;  // backward: re-derive the winners instead of reading codes
        float4 *         gradSorted; // backward: per sorted atom (dx, dy, dz, dr)
        const float4 *   molGrids;   // batches: per molecule (origin, 0), (spacing, 0); grid dx^3 each
    };

    // Query q and its molecule (batches sample one dx*dy*dz grid per molecule, molecule-major).
    __device__ __forceinline__ f3 query_point( const SdfArgs & s, uint64_t q, uint32_t & mol )
    {
        mol = 0;
        if ( s.points )
            return xyz( s.points[ q ] );
        const uint64_t per = uint64_t( s.dx ) * s.dy * s.dz;
        float          ox = s.ox, oy = s.oy, oz = s.oz, sx = s.sx, sy = s.sy, sz = s.sz;
        if ( s.molGrids )
        {
            mol              = uint32_t( q / per );
            const float4 o   = s.molGrids[ 2 * mol ], sp = s.molGrids[ 2 * mol + 1 ];
            ox = o.x, oy = o.y, oz = o.z, sx = sp.x, sy = sp.y, sz = sp.z;
            q -= uint64_t( mol ) * per;
        }
        const uint64_t ix = q % s.dx, iy = ( q / s.dx ) % s.dy, iz = q / ( uint64_t( s.dx ) * s.dy );
        return mk( ox + sx * float( ix ), oy + sy * float( iy ), oz + sz * float( iz ) );
    }

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

    // Evaluates the field at q; returns the value and the winner code.
    __device__ float sdf_eval( const Buffers & b, const SdfArgs & s, f3 q, uint32_t mol, uint64_t & code )
    {
        const float    rp   = b.probeRadius;
        const float    rmax = __uint_as_float( b.ctl->radiusMaxBits );
        const float    D    = rp + s.band;
        const Molecule M    = b.mols[ mol ];
        const uint32_t Iw   = b.ctl->intersectionWithNbNb;
        const uint32_t extC = s.exterior ? *s.extComp : 0u;
        const float    search = D + rmax + rp + 1e-3f;
        const int      reach  = max( int( ceilf( search / M.cell ) ), 1 );
        // nearest cell, clamped (queries may lie outside the brick)
        int gx, gy, gz;
        cell_of( M, q.x, q.y, q.z, gx, gy, gz );
        const int x0 = max( gx - reach, 0 ), x1 = min( gx + reach, M.dx - 1 );
        const int y0 = max( gy - reach, 0 ), y1 = min( gy + reach, M.dy - 1 );
        const int z0 = max( gz - reach, 0 ), z1 = min( gz + reach, M.dz - 1 );

        float    g = 3.0e38f, best = D, gExt = 3.0e38f;
        uint32_t gA = 0xffffffffu;
        uint64_t bc = code_of( SdfClamped, 0 );
        // pass 1: nearest sphere (outside distance), arcs and vertices
        for ( int zz = z0; zz <= z1; zz++ )
            for ( int yy = y0; yy <= y1; yy++ )
            {
                const uint32_t h0 = cell_index( M, x0, yy, zz );
                const uint32_t a0 = b.cellStart[ h0 ], a1 = b.cellStart[ h0 + uint32_t( x1 - x0 ) + 1 ];
                for ( uint32_t a = a0; a < a1; a++ )
                {
                    const float4 aa = b.atoms[ a ];
                    const float  sd = sqrtf( len2( q - xyz( aa ) ) ) - ( aa.w + rp );
                    if ( sd < g || ( sd == g && a < gA ) )
                        g = sd, gA = a;
                    if ( s.exterior && s.comp[ a ] == extC )
                        gExt = fminf( gExt, sd );
                    if ( sd > best || ( s.exterior && s.comp[ a ] != extC ) )
                        continue; // every element of a lies on sphere a
                    for ( uint32_t f = b.fullBase[ a ]; f < b.fullBase[ a + 1 ]; f++ )
                    {
                        const uint2 fc = b.fullCircles[ f ];
                        f3          ctr, nrm;
                        float       rad;
                        circle_world( b, fc.x, fc.y, ctr, nrm, rad );
                        const f3    v   = q - ctr;
                        const float ax  = dot( v, nrm );
                        const float rho = sqrtf( fmaxf( len2( v ) - ax * ax, 0.f ) );
                        const float d   = sqrtf( ax * ax + ( rho - rad ) * ( rho - rad ) );
                        if ( d < best )
                            best = d, bc = code_of( SdfCircle, f );
                    }
                    const uint32_t s0 = b.segStart[ b.trimBase[ a ] ], s1 = b.segStart[ b.trimBase[ a + 1 ] ];
                    for ( uint32_t sg = s0; sg < s1; sg++ )
                    {
                        const uint4 seg = b.segments[ sg ];
                        f3          ctr, nrm;
                        float       rad;
                        circle_world( b, seg.x, seg.y, ctr, nrm, rad );
                        const f3    v   = q - ctr;
                        const float ax  = dot( v, nrm );
                        const f3    w   = v - nrm * ax;
                        const float rho = sqrtf( len2( w ) );
                        if ( !( rho > 0.f ) )
                            continue;
                        const float d = sqrtf( ax * ax + ( rho - rad ) * ( rho - rad ) );
                        if ( d >= best )
                            continue;
                        const f3 A = xyz( b.concavePos[ seg.z ] ) - ctr, B = xyz( b.concavePos[ seg.w ] ) - ctr;
                        if ( angle_ccw( A, w, nrm ) <= angle_ccw( A, B, nrm ) )
                            best = d, bc = code_of( SdfSegment, sg );
                    }
                    for ( int bucket = 0; bucket < 2; bucket++ )
                    {
                        const uint32_t v0 = bucket == 0 ? b.wStart[ a ] : Iw + b.nStart[ a ];
                        const uint32_t v1 = bucket == 0 ? b.wStart[ a + 1 ] : Iw + b.nStart[ a + 1 ];
                        for ( uint32_t v = v0; v < v1; v++ )
                        {
                            const float d = sqrtf( len2( q - xyz( b.concavePos[ v ] ) ) );
                            if ( d < best )
                                best = d, bc = code_of( SdfVertex, v );
                        }
                    }
                }
            }
        // pass 2: faces closer than the best element so far
        for ( int zz = z0; zz <= z1; zz++ )
            for ( int yy = y0; yy <= y1; yy++ )
            {
                const uint32_t h0 = cell_index( M, x0, yy, zz );
                const uint32_t a0 = b.cellStart[ h0 ], a1 = b.cellStart[ h0 + uint32_t( x1 - x0 ) + 1 ];
                for ( uint32_t a = a0; a < a1; a++ )
                {
                    const float4 aa   = b.atoms[ a ];
                    const f3     dq   = q - xyz( aa );
                    const float  dist = sqrtf( len2( dq ) );
                    const float  Ra   = aa.w + rp;
                    const float  sd   = dist - Ra;
                    if ( !( fabsf( sd ) < best ) || !( dist > 0.f ) || ( s.exterior && s.comp[ a ] != extC ) )
                        continue;
                    const f3       p   = dq * ( Ra / dist ); // relative to c_a
                    const uint32_t na  = b.nbCount[ a ];
                    bool           exp = true;
                    for ( uint32_t l = 0; l < na && exp; l++ )
                    {
                        const float4 al = b.atoms[ b.nbIds[ size_t( a ) * b.K + l ] ];
                        const float  Rl = al.w + rp;
                        exp             = len2( p - ( xyz( al ) - xyz( aa ) ) ) - Rl * Rl >= 0.f;
                    }
                    if ( exp )
                        best = fabsf( sd ), bc = code_of( SdfFace, a );
                }
            }
        float f;
        if ( gA == 0xffffffffu )
        {
            code = code_of( SdfClamped, 0 );
            return s.band;
        }
        // outside A: exterior unless the nearest sphere of the exterior component is farther than
        // the nearest sphere by more than 1e-4 A (ties between components count as exterior)
        if ( g >= 0.f && ( !s.exterior || gExt <= g + 1e-4f ) )
        {
            f    = g;
            code = code_of( SdfOutside, gA );
        }
        else
        {
            f    = -best;
            code = bc;
        }
        const float v = rp + f;
        if ( v >= s.band || v <= -s.band )
        {
            code = code_of( SdfClamped, 0 );
            return v >= s.band ? s.band : -s.band;
        }
        return v;
    }

    __global__ void __launch_bounds__( 256 ) sdf_kernel( Buffers b, SdfArgs s )
    {
        for ( uint64_t q = uint64_t( blockIdx.x ) * blockDim.x + threadIdx.x; q < s.n; q += uint64_t( gridDim.x ) * blockDim.x )
        {
            uint64_t    code;
            uint32_t    mol;
            const f3    qp = query_point( s, q, mol );
            const float v  = sdf_eval( b, s, qp, mol, code );
            s.out[ q ]    = v;
            if ( s.codes )
                s.codes[ q ] = code;
        }
    }

    // nu for 1..3 sphere constraints: solve sum_m nu_m a_m = e in the least-squares sense
    // (normal equations, FP64).
    __device__ void solve_nu( int m, const d3 * a, d3 e, double * nu )
    {
        double G[ 3 ][ 3 ], r[ 3 ];
        for ( int u = 0; u < m; u++ )
        {
            r[ u ] = dot( a[ u ], e );
            for ( int v = 0; v < m; v++ )
                G[ u ][ v ] = dot( a[ u ], a[ v ] );
        }
        if ( m == 1 )
        {
            nu[ 0 ] = r[ 0 ] / G[ 0 ][ 0 ];
            return;
        }
        if ( m == 2 )
        {
            const double det = G[ 0 ][ 0 ] * G[ 1 ][ 1 ] - G[ 0 ][ 1 ] * G[ 1 ][ 0 ];
            nu[ 0 ]          = ( r[ 0 ] * G[ 1 ][ 1 ] - G[ 0 ][ 1 ] * r[ 1 ] ) / det;
            nu[ 1 ]          = ( G[ 0 ][ 0 ] * r[ 1 ] - G[ 1 ][ 0 ] * r[ 0 ] ) / det;
            return;
        }
        // 3x3: Cramer's rule on the Gram system
        auto det3 = []( double A[ 3 ][ 3 ] ) {
            return A[ 0 ][ 0 ] * ( A[ 1 ][ 1 ] * A[ 2 ][ 2 ] - A[ 1 ][ 2 ] * A[ 2 ][ 1 ] ) - A[ 0 ][ 1 ] * ( A[ 1 ][ 0 ] * A[ 2 ][ 2 ] - A[ 1 ][ 2 ] * A[ 2 ][ 0 ] ) +
                   A[ 0 ][ 2 ] * ( A[ 1 ][ 0 ] * A[ 2 ][ 1 ] - A[ 1 ][ 1 ] * A[ 2 ][ 0 ] );
        };
        const double d = det3( G );
        for ( int c = 0; c < 3; c++ )
        {
            double A[ 3 ][ 3 ];
            for ( int u = 0; u < 3; u++ )
                for ( int v = 0; v < 3; v++ )
                    A[ u ][ v ] = v == c ? r[ u ] : G[ u ][ v ];
            nu[ c ] = det3( A ) / d;
        }
    }

    __device__ __forceinline__ void grad_add( float4 * g, uint32_t a, d3 dc, double dr, float w )
    {
        atomicAdd( &g[ a ].x, float( w * dc.x ) );
        atomicAdd( &g[ a ].y, float( w * dc.y ) );
        atomicAdd( &g[ a ].z, float( w * dc.z ) );
        atomicAdd( &g[ a ].w, float( w * dr ) );
    }

    __global__ void __launch_bounds__( 256 ) sdf_backward_kernel( Buffers b, SdfArgs s )
    {
        const double rp = b.probeRadius;
        for ( uint64_t q = uint64_t( blockIdx.x ) * blockDim.x + threadIdx.x; q < s.n; q += uint64_t( gridDim.x ) * blockDim.x )
        {
            const float w = s.weights[ q ];
            if ( w == 0.f )
                continue;
            uint32_t mol;
            const f3 qf = query_point( s, q, mol );
            uint64_t code;
            if ( !s.recompute )
                code = s.codes[ q ]; // stored by the forward pass
            else
                (void)sdf_eval( b, s, qf, mol, code );
            const uint32_t type = uint32_t( code >> 40 ), idx = uint32_t( code & 0xffffffffffull );
            if ( type == SdfClamped )
                continue;
            const d3 Q = mkd( qf );
            auto     C = [ & ]( uint32_t a ) { return mkd( xyz( b.atoms[ a ] ) ); };
            auto     R = [ & ]( uint32_t a ) { return double( b.atoms[ a ].w ) + rp; };
            if ( type == SdfOutside )
            {
                const d3     v  = Q - C( idx );
                const double dn = sqrt( dot( v, v ) );
                grad_add( s.gradSorted, idx, v * ( -1.0 / dn ), -1.0, w );
                continue;
            }
            uint32_t at[ 3 ];
            int      m = 0;
            d3       p;
            if ( type == SdfFace )
            {
                at[ 0 ]       = idx;
                m             = 1;
                const d3 v    = Q - C( idx );
                p             = C( idx ) + v * ( R( idx ) / sqrt( dot( v, v ) ) );
            }
            else if ( type == SdfVertex )
            {
                const int4 ids = b.concaveIds[ idx ];
                at[ 0 ] = uint32_t( ids.x ), at[ 1 ] = uint32_t( ids.y ), at[ 2 ] = uint32_t( ids.z );
                m       = 3;
                p       = mkd( xyz( b.concavePos[ idx ] ) );
            }
            else
            {
                uint32_t i, j;
                if ( type == SdfCircle )
                    i = b.fullCircles[ idx ].x, j = b.fullCircles[ idx ].y;
                else
                    i = b.segments[ idx ].x, j = b.segments[ idx ].y;
                at[ 0 ] = i, at[ 1 ] = j;
                m       = 2;
                // nearest point of the circle (FP64)
                const d3     ci = C( i ), cj = C( j );
                const double Ri = R( i ), Rj = R( j );
                const d3     dj = cj - ci;
                const double d2 = dot( dj, dj );
                const d3     nrm = dj * ( 1.0 / sqrt( d2 ) );
                const d3     ctr = ci + dj * ( ( Ri * Ri - Rj * Rj + d2 ) / ( 2.0 * d2 ) );
                const double rad = sqrt( fmax( Ri * Ri - dot( ctr - ci, ctr - ci ), 0.0 ) );
                const d3     v   = Q - ctr;
                const d3     wv  = v - nrm * dot( v, nrm );
                p                = ctr + wv * ( rad / sqrt( dot( wv, wv ) ) );
            }
            const d3     e  = p - Q;
            const double D  = sqrt( dot( e, e ) );
            if ( !( D > 0.0 ) )
                continue;
            d3     av[ 3 ];
            double nu[ 3 ];
            for ( int u = 0; u < m; u++ )
                av[ u ] = p - C( at[ u ] );
            solve_nu( m, av, e * ( 1.0 / D ), nu );
            for ( int u = 0; u < m; u++ )
                grad_add( s.gradSorted, at[ u ], av[ u ] * ( -nu[ u ] ), -nu[ u ] * R( at[ u ] ), w );
        }
    }

    __global__ void gather_grad_kernel( const float4 * gradSorted, const uint32_t * s2i, uint32_t n, float4 * gradInput )
    {
        const uint32_t a = blockIdx.x * blockDim.x + threadIdx.x;
        if ( a < n )
            gradInput[ s2i[ a ] ] = gradSorted[ a ];
    }

    // ---- surface components (exterior variant): union-find over the arcs --------------------------
    // parent[] is read with agent-scope atomic loads: other XCDs hook through atomics and a plain
    // load could be served from this XCD's L2 with a stale value.
    __device__ __forceinline__ uint32_t uf_load( uint32_t * parent, uint32_t a )
    {
        return __hip_atomic_load( parent + a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT );
    }
    __device__ __forceinline__ uint32_t uf_find( uint32_t * parent, uint32_t a )
    {
        uint32_t p = uf_load( parent, a );
        while ( p != a )
        {
            a = p;
            p = uf_load( parent, a );
        }
        return a;
    }

    __global__ void comp_init_kernel( uint32_t * parent, uint32_t n, unsigned long long * best, const float4 * atoms )
    {
        const uint32_t a = blockIdx.x * blockDim.x + threadIdx.x;
        if ( a >= n )
            return;
        parent[ a ] = a;
        // key: largest x + r, ties to the lowest index; float made order-preserving as unsigned
        const float    v    = atoms[ a ].x + atoms[ a ].w;
        uint32_t       bits = __float_as_uint( v );
        bits                = ( bits & 0x80000000u ) ? ~bits : ( bits | 0x80000000u );
        atomicMax( best, ( (unsigned long long)bits << 32 ) | ( 0xffffffffu - a ) );
    }

    __global__ void comp_hook_kernel( Buffers b, uint32_t * parent, uint32_t * changed )
    {
        const uint32_t Cf = b.ctl->fullCircleNb, G = b.ctl->segmentNb;
        for ( uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < Cf + G; e += gridDim.x * blockDim.x )
        {
            uint32_t u, v;
            if ( e < Cf )
                u = b.fullCircles[ e ].x, v = b.fullCircles[ e ].y;
            else
                u = b.segments[ e - Cf ].x, v = b.segments[ e - Cf ].y;
            for ( ;; )
            {
                u = uf_find( parent, u );
                v = uf_find( parent, v );
                if ( u == v )
                    break;
                const uint32_t hi = max( u, v ), lo = min( u, v );
                if ( atomicCAS( parent + hi, hi, lo ) == hi )
                {
                    *changed = 1u;
                    break;
                }
            }
        }
    }

    __global__ void comp_compress_kernel( uint32_t * parent, uint32_t n, const unsigned long long * best, uint32_t * extComp )
    {
        const uint32_t a = blockIdx.x * blockDim.x + threadIdx.x;
        if ( a < n )
            parent[ a ] = uf_find( parent, a );
        if ( a == 0 )
        {
            const uint32_t top = 0xffffffffu - uint32_t( *best & 0xffffffffu );
            extComp[ 0 ]       = uf_find( parent, top );
        }
    }
} // namespace sesdf

using namespace sesdf;

namespace
{
    sesdf_status prepare_components( Handle & h )
    {
        if ( h.sdf.component && h.sdf.builtEpoch == h.hctl->epoch )
            return SESDF_OK;
        const uint32_t n    = h.n;
        const size_t   nPad = ( size_t( n ) + 2 + 1 ) & ~size_t( 1 ); // 8-byte aligned slot for the argmax key
        if ( !h.sdf.component )
        {
            void * p = nullptr;
            if ( hipMalloc( &p, nPad * 4 + 16 ) != hipSuccess )
                return SESDF_ERR_HIP;
            h.sdfAllocs.push_back( p );
            h.sdfBytes += nPad * 4 + 16;
            h.sdf.component = static_cast<uint32_t *>( p );
        }
        uint32_t *           parent  = h.sdf.component;
        uint32_t *           extComp = parent + n;
        uint32_t *           changed = parent + n + 1;
        unsigned long long * best    = reinterpret_cast<unsigned long long *>( parent + nPad );
        hipStream_t          s       = h.stream;
        // small fill of the reduction slot with a kernel-free path: host copy
        const unsigned long long zero64 = 0;
        if ( hipMemcpyAsync( best, &zero64, 8, hipMemcpyHostToDevice, s ) != hipSuccess )
            return SESDF_ERR_HIP;
        comp_init_kernel<<<( n + 255 ) / 256, 256, 0, s>>>( parent, n, best, h.b.atoms );
        for ( int it = 0; it < 64; it++ )
        {
            uint32_t one = 0, flag = 0;
            if ( hipMemcpyAsync( changed, &one, 4, hipMemcpyHostToDevice, s ) != hipSuccess )
                return SESDF_ERR_HIP;
            comp_hook_kernel<<<1024, 256, 0, s>>>( h.b, parent, changed );
            if ( hipMemcpyAsync( &flag, changed, 4, hipMemcpyDeviceToHost, s ) != hipSuccess || hipStreamSynchronize( s ) != hipSuccess )
                return SESDF_ERR_HIP;
            if ( !flag )
                break;
        }
        comp_compress_kernel<<<( n + 255 ) / 256, 256, 0, s>>>( parent, n, best, extComp );
        if ( hipGetLastError() != hipSuccess )
            return SESDF_ERR_HIP;
        h.sdf.builtEpoch = h.hctl->epoch;
        return SESDF_OK;
    }

    sesdf_status sdf_common( Handle & h, SdfArgs & s, int flags )
    {
        if ( !h.built )
            return SESDF_ERR_NOT_BUILT;
        // the field is defined on the complete surface; batches sample per-molecule grids only
        if ( h.mode != 0 || h.K > 320 || ( h.mols.size() != 1 && ( !s.molGrids || ( flags & 1 ) ) ) )
            return SESDF_ERR_UNSUPPORTED;
        if ( !( s.band > 0.f ) )
            return SESDF_ERR_INVALID;
        if ( hipSetDevice( h.device ) != hipSuccess )
            return SESDF_ERR_HIP;
        s.exterior = flags & 1;
        if ( s.exterior )
        {
            sesdf_status st = prepare_components( h );
            if ( st != SESDF_OK )
                return st;
            s.comp    = h.sdf.component;
            s.extComp = h.sdf.component + h.n;
        }
        return SESDF_OK;
    }

    uint32_t grid_blocks( uint64_t n )
    {
        const uint64_t want = ( n + 255 ) / 256;
        const uint64_t cap  = uint64_t( device_cu_count() ) * 32;
        return uint32_t( std::max<uint64_t>( 1, std::min( want, cap ) ) );
    }
} // namespace

extern "C"
{
    sesdf_status sesdf_sample_points_ex( sesdf_handle * hh, const float * d_points, uint64_t npoints, float band, int flags, float * d_out,
                                         uint64_t * d_codes )
    {
        if ( !hh || ( npoints && ( !d_points || !d_out ) ) )
            return SESDF_ERR_INVALID;
        Handle & h = *reinterpret_cast<Handle *>( hh );
        SdfArgs  s {};
        s.points     = reinterpret_cast<const float4 *>( d_points );
        s.n          = npoints;
        s.band       = band;
        s.out        = d_out;
        s.codes      = d_codes;
        sesdf_status st = sdf_common( h, s, flags );
        if ( st != SESDF_OK || npoints == 0 )
            return st;
        sdf_kernel<<<grid_blocks( npoints ), 256, 0, h.stream>>>( h.b, s );
        return hipGetLastError() == hipSuccess ? SESDF_OK : SESDF_ERR_HIP;
    }

    // Surface component label of every sorted atom (host_labels[atom_nb]) and the label of the
    // exterior component (component containing the atom with the largest x + r).
    sesdf_status sesdf_surface_components( sesdf_handle * hh, uint32_t * host_labels, uint32_t * exterior_label )
    {
        if ( !hh || !host_labels || !exterior_label )
            return SESDF_ERR_INVALID;
        Handle & h = *reinterpret_cast<Handle *>( hh );
        SdfArgs  s {};
        s.band          = 1.f;
        sesdf_status st = sdf_common( h, s, 1 );
        if ( st != SESDF_OK )
            return st;
        if ( hipMemcpyAsync( host_labels, h.sdf.component, size_t( h.n ) * 4, hipMemcpyDeviceToHost, h.stream ) != hipSuccess ||
             hipMemcpyAsync( exterior_label, h.sdf.component + h.n, 4, hipMemcpyDeviceToHost, h.stream ) != hipSuccess ||
             hipStreamSynchronize( h.stream ) != hipSuccess )
            return SESDF_ERR_HIP;
        return SESDF_OK;
    }

    sesdf_status sesdf_sample_points( sesdf_handle * hh, const float * d_points, uint64_t npoints, float band, float * d_out )
    {
        return sesdf_sample_points_ex( hh, d_points, npoints, band, 0, d_out, nullptr );
    }

    sesdf_status sesdf_sample_grid_ex( sesdf_handle * hh, const sesdf_grid_desc * grid, float band, int flags, float * d_out )
    {
        if ( !hh || !grid || !d_out )
            return SESDF_ERR_INVALID;
        Handle & h = *reinterpret_cast<Handle *>( hh );
        SdfArgs  s {};
        s.ox = grid->origin[ 0 ], s.oy = grid->origin[ 1 ], s.oz = grid->origin[ 2 ];
        s.sx = grid->spacing[ 0 ], s.sy = grid->spacing[ 1 ], s.sz = grid->spacing[ 2 ];
        s.dx = grid->dims[ 0 ], s.dy = grid->dims[ 1 ], s.dz = grid->dims[ 2 ];
        s.n    = uint64_t( s.dx ) * s.dy * s.dz;
        s.band = band;
        s.out  = d_out;
        sesdf_status st = sdf_common( h, s, flags );
        if ( st != SESDF_OK || s.n == 0 )
            return st;
        sdf_kernel<<<grid_blocks( s.n ), 256, 0, h.stream>>>( h.b, s );
        return hipGetLastError() == hipSuccess ? SESDF_OK : SESDF_ERR_HIP;
    }

    // Batches: one dim^3 vertex grid per molecule over its vdW bounding box grown by `margin`,
    // molecule-major into d_out (molecule_nb * dim^3 floats).
    sesdf_status sesdf_batch_sample_grids( sesdf_handle * hh, uint32_t dim, float margin, float band, float * d_out )
    {
        if ( !hh || !d_out || dim < 2 )
            return SESDF_ERR_INVALID;
        Handle &     h = *reinterpret_cast<Handle *>( hh );
        const size_t M = h.mols.size();
        if ( !h.sdf.molGrids || h.sdf.molGridDim != dim || h.sdf.molGridMargin != margin )
        {
            std::vector<float4> g( 2 * M );
            for ( size_t m = 0; m < M; m++ )
            {
                const float * bx = h.boxes.data() + 6 * m;
                g[ 2 * m ]       = make_float4( bx[ 0 ] - margin, bx[ 1 ] - margin, bx[ 2 ] - margin, 0.f );
                g[ 2 * m + 1 ]   = make_float4( ( bx[ 3 ] - bx[ 0 ] + 2 * margin ) / float( dim - 1 ), ( bx[ 4 ] - bx[ 1 ] + 2 * margin ) / float( dim - 1 ),
                                                ( bx[ 5 ] - bx[ 2 ] + 2 * margin ) / float( dim - 1 ), 0.f );
            }
            if ( !h.sdf.molGrids )
            {
                void * p = nullptr;
                if ( hipMalloc( &p, 2 * M * sizeof( float4 ) ) != hipSuccess )
                    return SESDF_ERR_HIP;
                h.sdfAllocs.push_back( p );
                h.sdfBytes += 2 * M * sizeof( float4 );
                h.sdf.molGrids = static_cast<float4 *>( p );
            }
            if ( hipMemcpy( h.sdf.molGrids, g.data(), 2 * M * sizeof( float4 ), hipMemcpyHostToDevice ) != hipSuccess )
                return SESDF_ERR_HIP;
            h.sdf.molGridDim    = dim;
            h.sdf.molGridMargin = margin;
        }
        SdfArgs s {};
        s.dx = s.dy = s.dz = dim;
        s.n        = uint64_t( dim ) * dim * dim * M;
        s.band     = band;
        s.out      = d_out;
        s.molGrids = h.sdf.molGrids;
        sesdf_status st = sdf_common( h, s, 0 );
        if ( st != SESDF_OK )
            return st;
        sdf_kernel<<<grid_blocks( s.n ), 256, 0, h.stream>>>( h.b, s );
        return hipGetLastError() == hipSuccess ? SESDF_OK : SESDF_ERR_HIP;
    }

    sesdf_status sesdf_sample_grid( sesdf_handle * hh, const sesdf_grid_desc * grid, float band, float * d_out )
    {
        return sesdf_sample_grid_ex( hh, grid, band, 0, d_out );
    }

    // d_codes: winner codes stored by sesdf_sample_points_ex (recompute = 0) or null (recompute = 1).
    sesdf_status sesdf_sdf_backward_ex( sesdf_handle * hh, const float * d_points, const float * d_weights, uint64_t npoints, float band,
                                        int flags, const uint64_t * d_codes, float * d_grad_sorted, float * d_grad )
    {
        if ( !hh || ( npoints && ( !d_points || !d_weights ) ) || !d_grad || !d_grad_sorted )
            return SESDF_ERR_INVALID;
        Handle & h = *reinterpret_cast<Handle *>( hh );
        SdfArgs  s {};
        s.points     = reinterpret_cast<const float4 *>( d_points );
        s.n          = npoints;
        s.band       = band;
        s.weights    = d_weights;
        s.codes      = const_cast<uint64_t *>( d_codes );
        s.recompute  = d_codes ? 0 : 1;
        s.gradSorted = reinterpret_cast<float4 *>( d_grad_sorted );
        sesdf_status st = sdf_common( h, s, flags );
        if ( st != SESDF_OK )
            return st;
        if ( hipMemsetAsync( d_grad_sorted, 0, size_t( h.n ) * 16, h.stream ) != hipSuccess )
            return SESDF_ERR_HIP;
        if ( npoints )
            sdf_backward_kernel<<<grid_blocks( npoints ), 256, 0, h.stream>>>( h.b, s );
        if ( h.n )
            gather_grad_kernel<<<( h.n + 255 ) / 256, 256, 0, h.stream>>>( s.gradSorted, h.b.sortedToInitial, h.n,
                                                                          reinterpret_cast<float4 *>( d_grad ) );
        return hipGetLastError() == hipSuccess ? SESDF_OK : SESDF_ERR_HIP;
    }

    sesdf_status sesdf_sdf_backward( sesdf_handle * hh, const float * d_points, const float * d_weights, uint64_t npoints, float band,
                                     int recompute, float * d_grad )
    {
        (void)recompute; // this form always recomputes the winners; see sesdf_sdf_backward_ex
        if ( !hh )
            return SESDF_ERR_INVALID;
        Handle & h = *reinterpret_cast<Handle *>( hh );
        if ( !h.sdf.arcs )
        {
            void * p = nullptr;
            if ( hipMalloc( &p, std::max<size_t>( size_t( h.n ) * 16, 16 ) ) != hipSuccess )
                return SESDF_ERR_HIP;
            h.sdfAllocs.push_back( p );
            h.sdfBytes += size_t( h.n ) * 16;
            h.sdf.arcs = static_cast<float4 *>( p ); // scratch: gradient in sorted order
        }
        return sesdf_sdf_backward_ex( hh, d_points, d_weights, npoints, band, 0, nullptr, reinterpret_cast<float *>( h.sdf.arcs ), d_grad );
    }
}
