// SDF sampling over the built complete surface, and its gradient with the topology held fixed.
//
// The reference ships no SDF evaluator (its renderer is not in the repository); this library
// defines  SDF(q) = clamp(rp + f(q), -band, band)  with f the signed distance to the boundary of
// the SAS union A (negative inside). Boundary elements are those of the complete-surface build:
//   faces    : sphere a where the radial projection of q is not inside another sphere of N(a);
//   arcs     : full circles and segments (A -> B counter-clockwise about n(i -> j));
//   vertices : concave patch (probe) positions.
// Outside A, f = min_a |q - c_a| - R_a. The field is 1-Lipschitz and its zero set is the SES.
// Exterior variant: only elements of the exterior boundary component count (components over arcs,
// vertices and faces, components.h), and points outside A in another region of the complement
// (probe-accessible cavities) are inside.
//
// Gradient (topology fixed): with p the nearest boundary point, e = (p - q) / |p - q| and nu
// solving sum_m nu_m (p - c_m) = e over the sphere constraints |p - c_m| = R_m of the winning
// element (1 for a face, 2 for an arc, 3 for a vertex):
//   dSDF/dc_m = -nu_m (p - c_m),  dSDF/dr_m = -nu_m R_m   (envelope theorem);
// outside A: dSDF/dc_a = -(q - c_a) / |q - c_a|, dSDF/dr_a = -1.

#include <algorithm>
#include <cstring>
#include <mutex>
#include <vector>

#include "handle.h"
#include "sort.h"

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

    // Flags of a compacted surface atom (high bits of its sorted index).
    constexpr uint32_t SurfOpen       = 1u << 31; // its face may be exposed
    constexpr uint32_t SurfHiddenCaps = 1u << 30; // a hidden circle's cap is not inside one visible cap
    constexpr uint32_t SurfAtom       = 0x3fffffffu;

    struct SdfArgs
    {
        const float4 *   points;     // n points (xyz_) or null: grid
        float            ox, oy, oz, sx, sy, sz;
        uint32_t         dx, dy, dz;
        uint64_t         n;
        float            band;
        int              exterior;
        Components       comp;       // surface components (exterior variant)
        float *          out;
        uint64_t *       codes;      // optional winner codes
        const float *    weights;    // backward
        int              recompute;  // backward: re-derive the winners instead of reading codes
        float4 *         gradSorted; // backward: per sorted atom (dx, dy, dz, dr)
        unsigned long long * gradFixed; // backward: 4 fixed-point accumulators per sorted atom; [4 n] max |w| bits
        const float4 *   molGrids;   // batches: per molecule (origin, 0), (spacing, 0); grid dx^3 each
        const uint32_t * perm;       // points given in cell order: result q goes to perm[q] (null: q)
        const float4 *   sAtoms;        // per sorted atom: centre, R = r + rp
        const float4 *   surfAtoms;     // surface atoms (compacted, cell order): centre, R
        const int4 *     surfRanges;    // per surface atom: (full circles, segments), (W vertices, N vertices)
        const uint32_t * surfIds;       // per surface atom: sorted index | SurfOpen | SurfHiddenCaps
        const uint32_t * surfCellStart; // cellNb + 1 offsets into the surface atoms
        const int2 *     surfCaps;      // per surface atom: range of its visible circles' caps
        const float4 *   caps;          // per visible circle (sector): axis, cos(half angle)
        const float4 *   segGeo;        // per segment: (centre, radius), (normal, arc > pi), n x A', B' x n
        int              empty;         // host side: no atoms, the field is +band everywhere
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

    // Point p of sphere a inside the SAS ball of a neighbour whose circle is hidden. Only atoms with
    // a hidden cap that no single visible cap contains are asked: their face may hold a region
    // that the visible caps leave open but hidden caps cover (a ring of visible caps around a hole
    // inside one hidden cap).
    __device__ __forceinline__ bool in_hidden_ball( const Buffers & b, uint32_t a, f3 p )
    {
        const uint32_t na = b.nbCount[ a ];
        bool           in = false;
        for ( uint32_t l = 0; l < na && !in; l++ )
        {
            if ( b.status[ size_t( a ) * b.K + l ] != 0 )
                continue;
            const float4 al = b.atoms[ b.nbIds[ size_t( a ) * b.K + l ] ];
            const float  Rl = al.w + b.probeRadius;
            in              = len2( p - xyz( al ) ) < Rl * Rl;
        }
        return in;
    }

    // In a wave-cooperative evaluation (COOP) all lanes walk the same cells and atoms, so a test
    // decides for the whole wave; otherwise every lane walks its own stencil.
    template<bool COOP>
    __device__ __forceinline__ bool any_lane( bool p )
    {
        if constexpr ( COOP )
            return ballot( p ) != 0;
        else
            return p;
    }
    template<bool COOP>
    __device__ __forceinline__ int uniform_int( int v )
    {
        if constexpr ( COOP )
            return __builtin_amdgcn_readfirstlane( v );
        else
            return v;
    }
    // Wave-wide min / max by DPP, within the VALU: pairs and quads (quad permutes), then rows
    // (shifts by 4 and 8), then row 15 / lane 31 broadcasts; the result gathers in lane 63 and is
    // returned from there (uniform). A lane whose DPP source lies outside the pattern keeps its
    // own value. Every lane of the wave must be active.
    template<int Ctrl>
    __device__ __forceinline__ int dpp_move( int v )
    {
        return __builtin_amdgcn_update_dpp( v, v, Ctrl, 0xf, 0xf, false );
    }
    template<class Op>
    __device__ __forceinline__ int wave_reduce( int v, Op op )
    {
        v = op( v, dpp_move<0xb1>( v ) );  // quad_perm [1, 0, 3, 2]
        v = op( v, dpp_move<0x4e>( v ) );  // quad_perm [2, 3, 0, 1]
        v = op( v, dpp_move<0x114>( v ) ); // row_shr:4
        v = op( v, dpp_move<0x118>( v ) ); // row_shr:8
        v = op( v, dpp_move<0x142>( v ) ); // row_bcast:15
        v = op( v, dpp_move<0x143>( v ) ); // row_bcast:31
        return __builtin_amdgcn_readlane( v, WaveSize - 1 );
    }
    __device__ __forceinline__ int wave_min_int( int v )
    {
        return wave_reduce( v, []( int a, int b ) { return min( a, b ); } );
    }
    __device__ __forceinline__ int wave_max_int( int v )
    {
        return wave_reduce( v, []( int a, int b ) { return max( a, b ); } );
    }
    __device__ __forceinline__ float wave_min_float( float v )
    {
        return __int_as_float( wave_reduce( __float_as_int( v ), []( int a, int b ) { return __float_as_int( fminf( __int_as_float( a ), __int_as_float( b ) ) ); } ) );
    }
    __device__ __forceinline__ float wave_max_float( float v )
    {
        return __int_as_float( wave_reduce( __float_as_int( v ), []( int a, int b ) { return __float_as_int( fmaxf( __int_as_float( a ), __int_as_float( b ) ) ); } ) );
    }

    // Bounding box of the wave's lanes that search surface elements (pass B).
    struct LaneBox
    {
        f3 lo, hi;
    };

    struct SdfLane
    {
        f3       q;
        bool     active;
        float    g, gExt, best;
        uint32_t gA;
        uint64_t bc;
    };

    // Cells within `reach` of the query cells, as row segments along x: first the block of query
    // cells grown by one (nearest atoms first, so that the bounds tighten early), then the rest.
    struct Stencil
    {
        int x0, x1, y0, y1, z0, z1;       // whole stencil
        int bx0, bx1, by0, by1, bz0, bz1; // near block
    };

    // Row segments of the stencil as fn(xa, xb, y, z, ord, dyz2): the near block first (ord = the
    // row's ordinal within the block), then the rest (ord = -1). dyz2 is the squared y/z gap from q
    // to the row, computed once per row for the caller's distance bound.
    template<class F>
    __device__ __forceinline__ void for_block( const Stencil & t, const Molecule & M, f3 q, F && fn )
    {
        int ord = 0;
        for ( int z = t.bz0; z <= t.bz1; z++ )
        {
            const float gz = slab_gap( M.oz + M.cell * float( z ), M.cell, q.z );
            for ( int y = t.by0; y <= t.by1; y++, ord++ )
            {
                const float gy = slab_gap( M.oy + M.cell * float( y ), M.cell, q.y );
                fn( t.bx0, t.bx1, y, z, ord, gz * gz + gy * gy );
            }
        }
    }
    template<class F>
    __device__ __forceinline__ void for_rest( const Stencil & t, const Molecule & M, f3 q, F && fn )
    {
        for ( int z = t.z0; z <= t.z1; z++ )
        {
            const float gz = slab_gap( M.oz + M.cell * float( z ), M.cell, q.z );
            for ( int y = t.y0; y <= t.y1; y++ )
            {
                const float gy   = slab_gap( M.oy + M.cell * float( y ), M.cell, q.y );
                const float dyz2 = gz * gz + gy * gy;
                if ( z >= t.bz0 && z <= t.bz1 && y >= t.by0 && y <= t.by1 )
                {
                    if ( t.x0 < t.bx0 )
                        fn( t.x0, t.bx0 - 1, y, z, -1, dyz2 );
                    if ( t.bx1 < t.x1 )
                        fn( t.bx1 + 1, t.x1, y, z, -1, dyz2 );
                }
                else
                    fn( t.x0, t.x1, y, z, -1, dyz2 );
            }
        }
    }

    // The stencil narrowed to the cells within r_lane of some lane that still searches (need):
    // rows outside it hold no atom whose row bound could pass. r_lane carries the same slack as the
    // row bound, plus 1e-3 A for the rounding of this cell range.
    template<bool COOP>
    __device__ __forceinline__ Stencil narrow( Stencil t, const Molecule & M, const SdfLane & L, bool need, float r )
    {
        constexpr int Big = 1 << 30;
        const float   rr  = r + 1e-3f;
        int x0 = need ? int( floorf( ( L.q.x - rr - M.ox ) / M.cell ) ) : Big, x1 = need ? int( floorf( ( L.q.x + rr - M.ox ) / M.cell ) ) : -Big;
        int y0 = need ? int( floorf( ( L.q.y - rr - M.oy ) / M.cell ) ) : Big, y1 = need ? int( floorf( ( L.q.y + rr - M.oy ) / M.cell ) ) : -Big;
        int z0 = need ? int( floorf( ( L.q.z - rr - M.oz ) / M.cell ) ) : Big, z1 = need ? int( floorf( ( L.q.z + rr - M.oz ) / M.cell ) ) : -Big;
        if constexpr ( COOP )
        {
            x0 = uniform_int<COOP>( wave_min_int( x0 ) ), x1 = uniform_int<COOP>( wave_max_int( x1 ) );
            y0 = uniform_int<COOP>( wave_min_int( y0 ) ), y1 = uniform_int<COOP>( wave_max_int( y1 ) );
            z0 = uniform_int<COOP>( wave_min_int( z0 ) ), z1 = uniform_int<COOP>( wave_max_int( z1 ) );
        }
        t.x0 = max( t.x0, x0 ), t.x1 = min( t.x1, x1 );
        t.y0 = max( t.y0, y0 ), t.y1 = min( t.y1, y1 );
        t.z0 = max( t.z0, z0 ), t.z1 = min( t.z1, z1 );
        return t;
    }

    // ---- primitive element tests -------------------------------------------------------------
    // (fsqrt / frcp: hardware forms, 1 ulp: the field is compared to FP64 at 1e-4 A, and the
    // correctly rounded forms cost over ten instructions each in these loops.)

    __device__ __forceinline__ void full_circle_test( const Buffers & b, SdfLane & L, uint32_t f )
    {
        const uint2 fc = b.fullCircles[ f ];
        f3          ctr, nrm;
        float       rad;
        circle_world( b, fc.x, fc.y, ctr, nrm, rad );
        const f3    v   = L.q - ctr;
        const float ax  = dot( v, nrm );
        const float rho = fsqrt( fmaxf( len2( v ) - ax * ax, 0.f ) );
        const float d   = fsqrt( ax * ax + ( rho - rad ) * ( rho - rad ) );
        if ( d < L.best )
            L.best = d, L.bc = code_of( SdfCircle, f );
    }

    // Segment geometry, 4 float4: (centre, radius), (normal n, 1 if the arc exceeds pi),
    // (n x A', 0), (B' x n, 0) with A', B' the endpoints relative to the centre. The foot w of q
    // in the circle plane lies on the arc from A' counter-clockwise to B' iff
    //   arc <= pi: w . (n x A') >= 0 and w . (B' x n) >= 0,
    //   arc >  pi: either holds (w is not strictly inside the complementary arc).
    __device__ __forceinline__ void segment_test( SdfLane & L, float4 g0, float4 g1, float4 g2, float4 g3, uint32_t sg )
    {
        const f3    nrm = xyz( g1 );
        const f3    v   = L.q - xyz( g0 );
        const float ax  = dot( v, nrm );
        const f3    w   = v - nrm * ax;
        const float rho = fsqrt( len2( w ) );
        const float d   = fsqrt( ax * ax + ( rho - g0.w ) * ( rho - g0.w ) );
        if ( !( rho > 0.f ) || !( d < L.best ) )
            return;
        const bool a = dot( w, xyz( g2 ) ) >= 0.f, bb = dot( w, xyz( g3 ) ) >= 0.f;
        if ( g1.w != 0.f ? ( a || bb ) : ( a && bb ) )
            L.best = d, L.bc = code_of( SdfSegment, sg );
    }

    __device__ __forceinline__ void vertex_test( SdfLane & L, float4 p, uint32_t v )
    {
        const float d = fsqrt( len2( L.q - xyz( p ) ) );
        if ( d < L.best )
            L.best = d, L.bc = code_of( SdfVertex, v );
    }

    // A face of sphere a can hold the nearest boundary point of a point q inside A only if q lies
    // inside ball a: were q outside it, the radial segment from q to the face point would leave
    // every ball through a boundary point nearer than the face point (the face point is outside
    // all other balls). Points of the exterior variant outside A (in cavities) may reach faces
    // from outside.
    template<bool EXT>
    __device__ __forceinline__ bool face_side( const SdfLane & L, float sd )
    {
        return sd < 0.f || ( EXT && L.g >= 0.f );
    }

    // ---- per-lane evaluation (scattered points): each lane walks its own cells ----------------
    template<bool EXT>
    __device__ __forceinline__ void lane_nearest( const SdfArgs & s, SdfLane & L, uint32_t a0, uint32_t a1, uint32_t extC )
    {
        for ( uint32_t a = a0; a < a1; a++ )
        {
            const float4 aa = s.sAtoms[ a ];
            const float  sd = fsqrt( len2( L.q - xyz( aa ) ) ) - aa.w;
            if ( sd < L.g || ( sd == L.g && a < L.gA ) )
                L.g = sd, L.gA = a;
            if ( EXT && s.comp.extStatus[ a ] == 1 )
                L.gExt = fminf( L.gExt, sd );
        }
    }

    // Arcs and vertices of the surface atoms [k0, k1) (each lies on its atom's sphere, at distance
    // >= |sd| from q), then their faces: the radial projection of q is exposed unless it lies in
    // the cap of a visible circle (hidden circles' caps lie inside visible ones).
    template<bool EXT>
    __device__ __forceinline__ void lane_row( const Buffers & b, const SdfArgs & s, SdfLane & L, uint32_t k0, uint32_t k1, uint32_t extC,
                                              uint32_t Iw, bool elems, bool faces )
    {
        for ( uint32_t k = k0; elems && k < k1; k++ )
        {
            const uint32_t a  = s.surfIds[ k ] & SurfAtom;
            const uint32_t st = EXT ? s.comp.extStatus[ a ] : 1u; // 2: element by element
            if ( st == 0 )
                continue;
            const float4 aa = s.surfAtoms[ k ];
            if ( !( fabsf( fsqrt( len2( L.q - xyz( aa ) ) ) - aa.w ) < L.best ) )
                continue;
            const int4 r0 = s.surfRanges[ 2 * k ], r1 = s.surfRanges[ 2 * k + 1 ];
            for ( int f = r0.x; f < r0.y; f++ )
                if ( st == 1 || s.comp.label[ f ] == extC )
                    full_circle_test( b, L, uint32_t( f ) );
            for ( int sg = r0.z; sg < r0.w; sg++ )
                if ( st == 1 || s.comp.label[ s.comp.capFull + uint32_t( sg ) ] == extC )
                    segment_test( L, s.segGeo[ 4 * sg ], s.segGeo[ 4 * sg + 1 ], s.segGeo[ 4 * sg + 2 ], s.segGeo[ 4 * sg + 3 ], uint32_t( sg ) );
            for ( int v = r1.x; v < r1.y; v++ )
                if ( st == 1 || s.comp.label[ vertex_node( s.comp, uint32_t( v ) ) ] == extC )
                    vertex_test( L, b.concavePos[ v ], uint32_t( v ) );
            for ( int v = int( Iw ) + r1.z; v < int( Iw ) + r1.w; v++ )
                if ( st == 1 || s.comp.label[ vertex_node( s.comp, uint32_t( v ) ) ] == extC )
                    vertex_test( L, b.concavePos[ v ], uint32_t( v ) );
        }
        for ( uint32_t k = k0; faces && k < k1; k++ )
        {
            const uint32_t id = s.surfIds[ k ], a = id & SurfAtom;
            const uint32_t st = EXT ? s.comp.extStatus[ a ] : 1u;
            if ( !( id & SurfOpen ) || st == 0 )
                continue;
            const float4 aa   = s.surfAtoms[ k ];
            const f3     dq   = L.q - xyz( aa );
            const float  dist = fsqrt( len2( dq ) );
            const float  sd   = dist - aa.w;
            if ( !( fabsf( sd ) < L.best ) || !face_side<EXT>( L, sd ) || !( dist > 0.f ) )
                continue;
            const f3   u       = dq * frcp( dist );
            const int2 cr      = s.surfCaps[ k ];
            bool       exposed = true;
            for ( int c = cr.x; c < cr.y && exposed; c++ )
            {
                const float4 cap = s.caps[ c ]; // axis, cos(half angle)
                exposed          = !( dot( u, xyz( cap ) ) > cap.w );
            }
            if ( exposed && ( id & SurfHiddenCaps ) )
                exposed = !in_hidden_ball( b, a, xyz( aa ) + u * aa.w );
            if ( exposed && st == 2 )
                exposed = label_at( b, s.comp, a, u ) == extC;
            if ( exposed )
                L.best = fabsf( sd ), L.bc = code_of( SdfFace, a );
        }
    }

    // ---- wave-cooperative evaluation: records are loaded 64 at a time, one per lane, and
    // broadcast with readlane, or staged in the wave's LDS area and read at a uniform address --
    __device__ __forceinline__ float    rl( float v, int j ) { return __int_as_float( __builtin_amdgcn_readlane( __float_as_int( v ), j ) ); }
    __device__ __forceinline__ int      rl( int v, int j ) { return __builtin_amdgcn_readlane( v, j ); }
    __device__ __forceinline__ uint32_t rl( uint32_t v, int j ) { return uint32_t( __builtin_amdgcn_readlane( int( v ), j ) ); }
    __device__ __forceinline__ float4   rl( float4 v, int j ) { return make_float4( rl( v.x, j ), rl( v.y, j ), rl( v.z, j ), rl( v.w, j ) ); }
    __device__ __forceinline__ int4     rl( int4 v, int j ) { return make_int4( rl( v.x, j ), rl( v.y, j ), rl( v.z, j ), rl( v.w, j ) ); }
    __device__ __forceinline__ float    uniform_float( float v ) { return __int_as_float( __builtin_amdgcn_readfirstlane( __float_as_int( v ) ) ); }
    // Value at a wave-uniform index, read through the constant address space (nothing the field
    // kernels read is written while they run): a scalar load into SGPRs. The chunk's records then
    // need not stay in VGPRs, one per lane, across the element and face loops, nor do the cell
    // ranges, the molecule and the build's control values. At a per-lane index it is an ordinary
    // vector load.
    template<class T>
    __device__ __forceinline__ T cload( const T * p, uint32_t i )
    {
#if defined( __HIP_DEVICE_COMPILE__ )
        using ConstT = const __attribute__( ( address_space( 4 ) ) ) T;
        return ( (ConstT *)( p ) )[ i ];
#else
        return p[ i ]; // host pass: never executed
#endif
    }

    // LDS stage of a wave (float4): a chunk's segments (4 each), then its W and N vertices, or the
    // caps of a chunk's faces. 4 KB per wave (with the masks, 16 KB per block of four) leaves room
    // for eight waves per SIMD; larger stages cost occupancy and ran slower.
    constexpr int StageSegs  = 40;
    constexpr int StageVerts = 32; // at most one per lane
    constexpr int StageW     = 4 * StageSegs, StageN = StageW + StageVerts;
    constexpr int StageSize  = StageN + StageVerts;
    static_assert( StageVerts <= WaveSize, "vertices are staged one per lane" );
    constexpr int MaskRows  = 32;  // near-block rows whose element-gate masks are kept (2 chunks each)

    template<bool EXT>
    __device__ __forceinline__ void coop_nearest( const SdfArgs & s, SdfLane & L, uint32_t a0, uint32_t a1, uint32_t extC )
    {
        const uint32_t lane = uint32_t( lane_id() );
        for ( uint32_t base = a0; base < a1; base += WaveSize )
        {
            const uint32_t a    = base + lane;
            const bool     ok   = a < a1;
            const float4   mine = ok ? s.sAtoms[ a ] : make_float4( 0.f, 0.f, 0.f, 0.f );
            const uint32_t ext  = ok && EXT && s.comp.extStatus[ a ] == 1 ? 1u : 0u;
            const int      cnt  = int( min( uint32_t( WaveSize ), a1 - base ) );
            for ( int j = 0; j < cnt; j++ )
            {
                const float4   aa = rl( mine, j );
                const float    sd = fsqrt( len2( L.q - xyz( aa ) ) ) - aa.w;
                const uint32_t ai = base + uint32_t( j );
                if ( sd < L.g || ( sd == L.g && ai < L.gA ) )
                    L.g = sd, L.gA = ai;
                if ( rl( ext, j ) )
                    L.gExt = fminf( L.gExt, sd );
            }
        }
    }

    // Surface atoms [k0, k1) of one row segment, 64 per chunk: arcs and vertices (elems), then faces
    // (faces). A face can win only where the element gate passed (|sd| < best, and best only
    // decreases), so faces are tested for the atoms of the chunk's element-gate mask: from the
    // element loop just run, or, for near-block rows (ord >= 0) whose elements ran in an earlier
    // pass, from the wave's mask store.
    template<bool EXT>
    __device__ __forceinline__ void coop_row( const Buffers & b, const SdfArgs & s, SdfLane & L, bool need, uint32_t k0, uint32_t k1, uint32_t extC,
                                              uint32_t Iw, bool elems, bool faces, int ord, float4 * stage, uint64_t * masks,
                                              const LaneBox & box )
    {
        const uint32_t lane = uint32_t( lane_id() );
        for ( uint32_t base = k0; base < k1; base += WaveSize )
        {
            const uint32_t chunk    = ( base - k0 ) / WaveSize;
            const bool     keepMask = ord >= 0 && ord < MaskRows && chunk < 2;
            uint64_t       passMask = 0;
            const uint32_t k      = base + lane;
            const bool     ok     = k < k1;
            const float4   myAtom = ok ? s.surfAtoms[ k ] : make_float4( 0.f, 0.f, 0.f, 0.f );
            const uint32_t myId   = ok ? s.surfIds[ k ] : 0u;
            const uint32_t mySt   = !ok ? 0u : EXT ? s.comp.extStatus[ myId & SurfAtom ] : 1u;
            const uint32_t myExt  = mySt != 0 ? 1u : 0u;
            const uint64_t chkMask = ballot( mySt == 2 ); // atoms whose elements are checked one by one
            const int      cnt    = int( min( uint32_t( WaveSize ), k1 - base ) );
            if ( elems )
            {
                // Chunk prefilter, one atom per lane: the shell |d - R| < best of some searching lane
                // must meet the box of those lanes (largest best, nearest / farthest box point).
                const float bestMax = uniform_float( wave_max_float( need ? L.best : 0.f ) );
                const f3    c       = xyz( myAtom );
                const float nx = fmaxf( fmaxf( box.lo.x - c.x, c.x - box.hi.x ), 0.f ), ny = fmaxf( fmaxf( box.lo.y - c.y, c.y - box.hi.y ), 0.f ),
                            nz = fmaxf( fmaxf( box.lo.z - c.z, c.z - box.hi.z ), 0.f );
                const float fx = fmaxf( fabsf( c.x - box.lo.x ), fabsf( c.x - box.hi.x ) ), fy = fmaxf( fabsf( c.y - box.lo.y ), fabsf( c.y - box.hi.y ) ),
                            fz = fmaxf( fabsf( c.z - box.lo.z ), fabsf( c.z - box.hi.z ) );
                const float dNear = fsqrt( nx * nx + ny * ny + nz * nz ), dFar = fsqrt( fx * fx + fy * fy + fz * fz );
                // 1e-4 A of slack covers the rounding of the two bounds
                uint64_t pre = ballot( ok && myExt && dNear - myAtom.w < bestMax + 1e-4f && myAtom.w - dFar < bestMax + 1e-4f );
                // the chunk's atoms are consecutive in sorted order, so the records they own are
                // contiguous; staged in LDS (up to 64 of each kind) on the first atom that needs them
                const int4 f0 = cload( s.surfRanges, 2 * base ), f1 = cload( s.surfRanges, 2 * base + 1 );
                const int4 l0 = cload( s.surfRanges, 2 * ( base + cnt - 1 ) ), l1 = cload( s.surfRanges, 2 * ( base + cnt - 1 ) + 1 );
                const int  s0 = f0.z, s1 = l0.w, w0 = f1.x, w1 = l1.y, n0 = int( Iw ) + f1.z, n1 = int( Iw ) + l1.w;
                const bool segIn = s1 - s0 <= StageSegs, wIn = w1 - w0 <= StageVerts, nIn = n1 - n0 <= StageVerts;
                bool       staged = false;
                while ( pre )
                {
                    const int j = __ffsll( (unsigned long long)pre ) - 1;
                    pre &= pre - 1;
                    const float4 aa = cload( s.surfAtoms, base + j );
                    if ( !any_lane<true>( need && fabsf( fsqrt( len2( L.q - xyz( aa ) ) ) - aa.w ) < L.best ) )
                        continue;
                    passMask |= uint64_t( 1 ) << j;
                    if ( !staged )
                    {
                        wave_sync();
                        for ( int i = int( lane ); segIn && i < 4 * ( s1 - s0 ); i += WaveSize )
                            stage[ i ] = s.segGeo[ 4 * s0 + i ];
                        if ( wIn && int( lane ) < w1 - w0 )
                            stage[ StageW + lane ] = b.concavePos[ w0 + int( lane ) ];
                        if ( nIn && int( lane ) < n1 - n0 )
                            stage[ StageN + lane ] = b.concavePos[ n0 + int( lane ) ];
                        wave_sync();
                        staged = true;
                    }
                    const int4 r0 = cload( s.surfRanges, 2 * ( base + j ) ), r1 = cload( s.surfRanges, 2 * ( base + j ) + 1 );
                    const bool chk = ( chkMask >> j ) & 1;
                    for ( int f = r0.x; f < r0.y; f++ )
                        if ( !chk || s.comp.label[ f ] == extC )
                            full_circle_test( b, L, uint32_t( f ) );
                    for ( int sg = r0.z; sg < r0.w; sg++ )
                    {
                        if ( chk && s.comp.label[ s.comp.capFull + uint32_t( sg ) ] != extC )
                            continue;
                        const float4 * g = segIn ? stage + 4 * ( sg - s0 ) : s.segGeo + 4 * sg;
                        segment_test( L, g[ 0 ], g[ 1 ], g[ 2 ], g[ 3 ], uint32_t( sg ) );
                    }
                    for ( int v = r1.x; v < r1.y; v++ )
                        if ( !chk || s.comp.label[ vertex_node( s.comp, uint32_t( v ) ) ] == extC )
                            vertex_test( L, wIn ? stage[ StageW + v - w0 ] : b.concavePos[ v ], uint32_t( v ) );
                    for ( int v = int( Iw ) + r1.z; v < int( Iw ) + r1.w; v++ )
                        if ( !chk || s.comp.label[ vertex_node( s.comp, uint32_t( v ) ) ] == extC )
                            vertex_test( L, nIn ? stage[ StageN + v - n0 ] : b.concavePos[ v ], uint32_t( v ) );
                }
            }
            if ( elems && !faces && keepMask )
            {
                if ( lane == 0 )
                    masks[ 2 * ord + chunk ] = passMask;
                wave_sync();
            }
            if ( faces )
            {
                uint64_t cand = ~uint64_t( 0 );
                if ( elems )
                    cand = passMask;
                else if ( keepMask )
                {
                    const uint64_t m = masks[ 2 * ord + chunk ];
                    cand             = uint64_t( uint32_t( __builtin_amdgcn_readfirstlane( int( uint32_t( m ) ) ) ) ) |
                           ( uint64_t( uint32_t( __builtin_amdgcn_readfirstlane( int( uint32_t( m >> 32 ) ) ) ) ) << 32 );
                }
                cand &= ballot( ok && ( myId & SurfOpen ) && myExt );
                const int  c0 = cload( s.surfCaps, base ).x, c1 = cload( s.surfCaps, base + cnt - 1 ).y;
                const bool capIn  = c1 - c0 <= StageSize;
                bool       staged = false;
                while ( cand )
                {
                    const int j = __ffsll( (unsigned long long)cand ) - 1;
                    cand &= cand - 1;
                    const uint32_t id      = cload( s.surfIds, base + j );
                    const float4   aa      = cload( s.surfAtoms, base + j );
                    const f3       dq      = L.q - xyz( aa );
                    const float    dist    = fsqrt( len2( dq ) );
                    const float    sd      = dist - aa.w;
                    bool           exposed = need && fabsf( sd ) < L.best && face_side<EXT>( L, sd ) && dist > 0.f;
                    if ( !any_lane<true>( exposed ) )
                        continue;
                    if ( capIn && !staged )
                    {
                        wave_sync();
                        for ( int i = int( lane ); i < c1 - c0; i += WaveSize )
                            stage[ i ] = s.caps[ c0 + i ];
                        wave_sync();
                        staged = true;
                    }
                    const f3  u  = dq * frcp( dist );
                    const int2 cr = cload( s.surfCaps, base + j );
                    const int  a0 = cr.x, a1 = cr.y;
                    // caps in groups of four (largest first): one exit test per group
                    for ( int c = a0; c < a1 && any_lane<true>( exposed ); c += 4 )
                    {
#pragma unroll
                        for ( int e = 0; e < 4; e++ )
                            if ( c + e < a1 )
                            {
                                const float4 cap = capIn ? stage[ c + e - c0 ] : s.caps[ c + e ]; // axis, cos(half angle)
                                exposed          = exposed && !( dot( u, xyz( cap ) ) > cap.w );
                            }
                    }
                    if ( ( id & SurfHiddenCaps ) && any_lane<true>( exposed ) )
                        exposed = exposed && !in_hidden_ball( b, id & SurfAtom, xyz( aa ) + u * aa.w );
                    if ( exposed && ( ( chkMask >> j ) & 1 ) )
                        exposed = label_at( b, s.comp, id & SurfAtom, u ) == extC;
                    if ( exposed )
                        L.best = fabsf( sd ), L.bc = code_of( SdfFace, id & SurfAtom );
                }
            }
        }
    }

    // Field value and winner code at L.q (molecule `mol`, identical in every lane when COOP).
    //   pass A (all atoms): nearest sphere g; with g >= rp + band nothing lies within the band and
    //     the value is +band; outside A (g >= 0, exterior component for the exterior variant) the
    //     value is rp + g;
    //   pass B (surface atoms only, inside points): nearest arc, vertex or exposed face; in the
    //     near block arcs and vertices come first, so that the costlier face tests are pruned by
    //     their distance; further rows are nearly always culled or cheap.
    // A row segment is skipped when the distance bound of its box cannot change the outcome.
    template<bool COOP, bool EXT>
    __device__ float sdf_eval_lanes( const Buffers & b, const SdfArgs & s, uint32_t mol, SdfLane & L, uint64_t & code, float4 * stage,
                                     uint64_t * masks )
    {
        const float    rp    = b.probeRadius;
        const float    Rmax  = __uint_as_float( cload( &b.ctl->radiusMaxBits, 0 ) ) + rp;
        const float    D     = rp + s.band;
        const Molecule M     = cload( b.mols, mol );
        const uint32_t Iw    = cload( &b.ctl->intersectionWithNbNb, 0 );
        const uint32_t extC  = EXT ? cload( s.comp.extComp, 0 ) : 0u;
        const int      reach = max( int( ceilf( ( D + Rmax + 1e-3f ) / M.cell ) ), 1 );
        int            gx, gy, gz;
        cell_of( M, L.q.x, L.q.y, L.q.z, gx, gy, gz );
        int cx0 = gx, cx1 = gx, cy0 = gy, cy1 = gy, cz0 = gz, cz1 = gz;
        if constexpr ( COOP )
        {
            constexpr int Big = 1 << 30;
            cx0 = uniform_int<COOP>( wave_min_int( L.active ? gx : Big ) ), cx1 = uniform_int<COOP>( wave_max_int( L.active ? gx : -Big ) );
            cy0 = uniform_int<COOP>( wave_min_int( L.active ? gy : Big ) ), cy1 = uniform_int<COOP>( wave_max_int( L.active ? gy : -Big ) );
            cz0 = uniform_int<COOP>( wave_min_int( L.active ? gz : Big ) ), cz1 = uniform_int<COOP>( wave_max_int( L.active ? gz : -Big ) );
        }
        Stencil t;
        t.x0 = max( cx0 - reach, 0 ), t.x1 = min( cx1 + reach, M.dx - 1 );
        t.y0 = max( cy0 - reach, 0 ), t.y1 = min( cy1 + reach, M.dy - 1 );
        t.z0 = max( cz0 - reach, 0 ), t.z1 = min( cz1 + reach, M.dz - 1 );
        t.bx0 = max( cx0 - 1, 0 ), t.bx1 = min( cx1 + 1, M.dx - 1 );
        t.by0 = max( cy0 - 1, 0 ), t.by1 = min( cy1 + 1, M.dy - 1 );
        t.bz0 = max( cz0 - 1, 0 ), t.bz1 = min( cz1 + 1, M.dz - 1 );
        // lower bound of |q - c_a| - R_a over a row segment [xa, xb] (1e-3 A covers the rounding of
        // the cell assignment)
        auto lower = [ & ]( int xa, int xb, float dyz2 ) {
            const float bx0 = M.ox + M.cell * float( xa ), bx1 = M.ox + M.cell * float( xb + 1 );
            const float gx  = fmaxf( fmaxf( bx0 - L.q.x, L.q.x - bx1 ), 0.f );
            return fsqrt( gx * gx + dyz2 ) - Rmax - 1e-3f;
        };

        // pass A. Once g < 0 the point is inside and g no longer matters. Otherwise a segment can
        // change the outcome only below g + 1e-4 (exterior test) or, for the complete variant,
        // below min(g, max(band - rp, 0)) (above, the value is clamped to +band).
        L.g = D, L.gExt = 3.0e38f, L.gA = 0xffffffffu;
        auto nearest = [ & ]( int xa, int xb, int y, int z, int, float dyz2 ) {
            if ( !any_lane<COOP>( L.active && L.g >= 0.f ) )
                return; // every point is inside: g no longer matters
            const float lb   = lower( xa, xb, dyz2 );
            const float gCut = EXT ? L.g + 1e-4f : fminf( L.g, fmaxf( s.band - rp, 0.f ) );
            if ( !any_lane<COOP>( L.active && L.g >= 0.f && lb <= gCut ) )
                return;
            const uint32_t h0 = cell_index( M, xa, y, z );
            const uint32_t a0 = cload( b.cellStart, h0 ), a1 = cload( b.cellStart, h0 + uint32_t( xb - xa ) + 1 );
            if constexpr ( COOP )
                coop_nearest<EXT>( s, L, a0, a1, extC );
            else
                lane_nearest<EXT>( s, L, a0, a1, extC );
        };
        for_block( t, M, L.q, nearest );
        {
            // the rest of the stencil, narrowed to what the bounds found in the near block allow
            const bool  needA = L.active && L.g >= 0.f;
            const float gCut  = EXT ? L.g + 1e-4f : fminf( L.g, fmaxf( s.band - rp, 0.f ) );
            for_rest( narrow<COOP>( t, M, L, needA, gCut + Rmax + 1e-3f ), M, L.q, nearest );
        }

        // 0: +band (nothing within the band), 1: outside A, 2: inside (pass B)
        int state = 0;
        if ( L.gA == 0xffffffffu || !( L.g < D ) )
        {
            // farther than rp + band from A: +band in the unbounded region, -band in a cavity (pass B
            // then finds no element within the band)
            if ( EXT && L.active && far_point_in_cavity( b, s.comp, L.q, D, extC ) )
                state = 2;
        }
        else
        {
            // outside A the point lies in the region of the complement bounded by the face that holds
            // its projection onto the nearest sphere; ties within 1e-4 A count as exterior
            bool outside = L.g >= 0.f && ( !EXT || L.gExt <= L.g + 1e-4f );
            if ( EXT && L.g >= 0.f && !outside && s.comp.extStatus[ L.gA ] == 2 )
            {
                const f3    dq = L.q - xyz( s.sAtoms[ L.gA ] );
                const float dl = fsqrt( len2( dq ) );
                outside        = dl > 0.f && label_at( b, s.comp, L.gA, dq * ( 1.f / dl ) ) == extC;
            }
            state = outside ? 1 : 2;
        }
        const bool needB = L.active && state == 2;
        L.best = D, L.bc = code_of( SdfClamped, 0 );
        if ( any_lane<COOP>( needB ) )
        {
            LaneBox box {};
            if constexpr ( COOP )
            {
                box.lo = mk( uniform_float( wave_min_float( needB ? L.q.x : 3.0e38f ) ), uniform_float( wave_min_float( needB ? L.q.y : 3.0e38f ) ),
                             uniform_float( wave_min_float( needB ? L.q.z : 3.0e38f ) ) );
                box.hi = mk( uniform_float( wave_max_float( needB ? L.q.x : -3.0e38f ) ), uniform_float( wave_max_float( needB ? L.q.y : -3.0e38f ) ),
                             uniform_float( wave_max_float( needB ? L.q.z : -3.0e38f ) ) );
            }
            auto surface = [ & ]( int xa, int xb, int y, int z, int ord, float dyz2, bool elems, bool faces ) {
                if ( !any_lane<COOP>( needB && lower( xa, xb, dyz2 ) < L.best ) )
                    return;
                const uint32_t h0 = cell_index( M, xa, y, z );
                const uint32_t k0 = cload( s.surfCellStart, h0 ), k1 = cload( s.surfCellStart, h0 + uint32_t( xb - xa ) + 1 );
                // faces of a near-block row whose atoms all failed the element gate: nothing to test
                if constexpr ( COOP )
                    if ( !elems && ord >= 0 && ord < MaskRows && k1 - k0 <= 2 * WaveSize && ( masks[ 2 * ord ] | masks[ 2 * ord + 1 ] ) == 0 )
                        return;
                if constexpr ( COOP )
                    coop_row<EXT>( b, s, L, needB, k0, k1, extC, Iw, elems, faces, ord, stage, masks, box );
                else
                    lane_row<EXT>( b, s, L, k0, k1, extC, Iw, elems, faces );
            };
            for_block( t, M, L.q, [ & ]( int xa, int xb, int y, int z, int ord, float dyz2 ) { surface( xa, xb, y, z, ord, dyz2, true, false ); } );
            for_block( t, M, L.q, [ & ]( int xa, int xb, int y, int z, int ord, float dyz2 ) { surface( xa, xb, y, z, ord, dyz2, false, true ); } );
            for_rest( narrow<COOP>( t, M, L, needB, L.best + Rmax + 1e-3f ), M, L.q,
                      [ & ]( int xa, int xb, int y, int z, int ord, float dyz2 ) { surface( xa, xb, y, z, ord, dyz2, true, true ); } );
        }

        if ( state == 0 )
        {
            code = code_of( SdfClamped, 0 );
            return s.band;
        }
        const float v = rp + ( state == 1 ? L.g : -L.best );
        code          = state == 1 ? code_of( SdfOutside, L.gA ) : L.bc;
        if ( v >= s.band || v <= -s.band )
        {
            code = code_of( SdfClamped, 0 );
            return v >= s.band ? s.band : -s.band;
        }
        return v;
    }

    template<bool EXT>
    __device__ float sdf_eval( const Buffers & b, const SdfArgs & s, f3 q, uint32_t mol, uint64_t & code )
    {
        SdfLane L;
        L.q      = q;
        L.active = true;
        return sdf_eval_lanes<false, EXT>( b, s, mol, L, code, nullptr, nullptr );
    }

    // Cap of circle (a, slot l) on sphere a in FP64 from the FP32 inputs: pole (unit, towards the
    // neighbour), cos and sin of the half angle. False when the spheres meet in no circle.
    __device__ __forceinline__ bool cap64( const Buffers & b, d3 ca, double Ra, uint32_t a, uint32_t l, d3 & pole, double & cs, double & sn )
    {
        const float4 al = b.atoms[ b.nbIds[ size_t( a ) * b.K + l ] ];
        const double Rl = double( al.w ) + double( b.probeRadius );
        const d3     v  = mkd( xyz( al ) ) - ca;
        const double d2 = dot( v, v ), d = sqrt( d2 );
        if ( !( d > fabs( Ra - Rl ) && d < Ra + Rl ) )
            return false;
        const double t = ( Ra * Ra - Rl * Rl + d2 ) / ( 2.0 * d );
        pole           = v * ( 1.0 / d );
        cs             = t / Ra;
        sn             = sqrt( fmax( Ra * Ra - t * t, 0.0 ) ) / Ra;
        return true;
    }

    // Containment of a hidden circle's cap (pole ph, cos ch, sin sh) in a visible cap (pole pv,
    // cos cv), FP64: the angle between the poles plus the hidden half angle is at most the visible
    // half angle (cosine and sine of the sum).
    __device__ __forceinline__ bool cap_in_cap( d3 ph, double ch, double sh, d3 pv, double cv )
    {
        const double c    = dot( ph, pv );
        const d3     x    = cross( ph, pv );
        const double sn   = sqrt( dot( x, x ) );
        const double cSum = c * ch - sn * sh, sSum = sn * ch + c * sh;
        return sSum >= 0.0 && cSum >= cv + 1e-12;
    }

    // Per sorted atom: centre and R; surface flag (owns an arc or vertex, or its face may be
    // exposed) for the compaction scan. A face can be exposed only if the atom has a visible
    // circle, or no neighbour cuts a cap off its sphere: otherwise the exposed part would be
    // bounded by arcs of visible circles. info: face may be exposed (bit 0); some hidden circle's
    // cap lies in no single visible cap, so the face test also asks the hidden neighbours' balls
    // (bit 1; almost every hidden cap lies inside one visible cap).
    // One wave per atom, lanes over its neighbours. The hidden-cap test runs only for atoms whose
    // face may be exposed (bit 1 means nothing otherwise): hidden circles 64 at a time, one per
    // lane, against the visible circles' caps, computed 64 at a time into LDS, so a (hidden,
    // visible) pair costs a broadcast read and no memory access.
    struct CapF64
    {
        double x, y, z, c;
    };
    __global__ void __launch_bounds__( 256 ) sdf_pack_kernel( Buffers b, float4 * pa, uint32_t * surfFlag, uint32_t * info )
    {
        __shared__ CapF64 visAll[ 256 / WaveSize ][ WaveSize ];
        CapF64 *          vis  = visAll[ wave_in_block() ];
        const uint32_t    lane = uint32_t( lane_id() );
        const uint32_t    a    = xcd_block() * ( 256 / WaveSize ) + uint32_t( wave_in_block() );
        if ( a >= b.n )
            return;
        // inside: the sphere lies in a neighbour's ball; cap: some neighbour robustly cuts a
        // non-empty cap (the caps of visible circles then bound any exposed part)
        const float4   at     = b.atoms[ a ];
        const float    Ra     = at.w + b.probeRadius;
        const uint32_t na     = b.nbCount[ a ];
        bool           inside = false, cap = false;
        for ( uint32_t l = lane; l < na; l += WaveSize )
        {
            const float4 al = b.atoms[ b.nbIds[ size_t( a ) * b.K + l ] ];
            const float  Rl = al.w + b.probeRadius;
            const float  d  = sqrtf( len2( xyz( al ) - xyz( at ) ) );
            inside          = inside || d + Ra <= Rl;
            cap             = cap || ( d < Ra + Rl - 1e-3f && d + Rl > Ra + 1e-3f );
        }
        inside           = any_lane<true>( inside );
        cap              = any_lane<true>( cap );
        const bool open  = !inside && ( b.vStart[ a + 1 ] > b.vStart[ a ] || !cap );
        bool       hidden = false;
        if ( open )
        {
            const d3     ca = mkd( xyz( at ) );
            const double Rd = double( at.w ) + double( b.probeRadius );
            for ( uint32_t hb = 0; hb < na && !hidden; hb += WaveSize )
            {
                // this lane's hidden circle stays "open" until a visible cap contains its cap
                const uint32_t h = hb + lane;
                d3             ph {};
                double         ch = 0.0, sh = 0.0;
                bool           openCap = h < na && b.status[ size_t( a ) * b.K + h ] == 0 && cap64( b, ca, Rd, a, h, ph, ch, sh );
                for ( uint32_t vb = 0; vb < na && any_lane<true>( openCap ); vb += WaveSize )
                {
                    const uint32_t l = vb + lane;
                    d3             pv {};
                    double         cv = 0.0, sv = 0.0;
                    const bool     ok = l < na && b.status[ size_t( a ) * b.K + l ] != 0 && cap64( b, ca, Rd, a, l, pv, cv, sv );
                    wave_sync();
                    if ( ok )
                        vis[ lane ] = { pv.x, pv.y, pv.z, cv };
                    wave_sync();
                    for ( uint64_t vm = ballot( ok ); vm && any_lane<true>( openCap ); vm &= vm - 1 )
                    {
                        const CapF64 v = vis[ __ffsll( (unsigned long long)vm ) - 1 ];
                        if ( openCap && cap_in_cap( ph, ch, sh, { v.x, v.y, v.z }, v.c ) )
                            openCap = false;
                    }
                }
                hidden = any_lane<true>( openCap );
            }
        }
        if ( lane != 0 )
            return;
        pa[ a ]          = make_float4( at.x, at.y, at.z, Ra );
        const bool elems = b.fullBase[ a + 1 ] > b.fullBase[ a ] || b.segStart[ b.trimBase[ a + 1 ] ] > b.segStart[ b.trimBase[ a ] ] ||
                           b.wStart[ a + 1 ] > b.wStart[ a ] || b.nStart[ a + 1 ] > b.nStart[ a ];
        surfFlag[ a ]    = ( elems || open ) ? 1u : 0u;
        info[ a ]        = ( open ? 1u : 0u ) | ( open && hidden ? 2u : 0u );
    }

    // Compacted surface atoms in cell order and their per-cell offsets.
    __global__ void sdf_surface_kernel( Buffers b, const float4 * pa, const uint32_t * surfFlag, const uint32_t * info, const uint32_t * prefix,
                                        float4 * sa, int4 * sr, uint32_t * sid, int2 * sc, uint32_t * sCellStart )
    {
        const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
        if ( x <= b.cellNb )
            sCellStart[ x ] = prefix[ b.cellStart[ x ] ];
        if ( x >= b.n || !surfFlag[ x ] )
            return;
        const uint32_t a = x, k = prefix[ a ];
        sa[ k ]          = pa[ a ];
        sr[ 2 * k ]      = make_int4( int( b.fullBase[ a ] ), int( b.fullBase[ a + 1 ] ), int( b.segStart[ b.trimBase[ a ] ] ),
                                      int( b.segStart[ b.trimBase[ a + 1 ] ] ) );
        sr[ 2 * k + 1 ]  = make_int4( int( b.wStart[ a ] ), int( b.wStart[ a + 1 ] ), int( b.nStart[ a ] ), int( b.nStart[ a + 1 ] ) );
        sid[ k ]         = a | ( ( info[ a ] & 1u ) ? SurfOpen : 0u ) | ( ( info[ a ] & 2u ) ? SurfHiddenCaps : 0u );
        sc[ k ]          = make_int2( int( b.vStart[ a ] ), int( b.vStart[ a + 1 ] ) );
    }

    // Cap of every visible circle (sector: axis, half angle) and the geometry of every segment.
    __global__ void __launch_bounds__( 256 ) sdf_geometry_kernel( Buffers b, float4 * caps, float4 * segGeo )
    {
        const uint32_t nCaps = min( b.ctl->sectorNb, b.capSectors ), nSeg = min( b.ctl->segmentNb, b.capSegments );
        for ( uint32_t x = blockIdx.x * blockDim.x + threadIdx.x; x < max( nCaps, nSeg ); x += gridDim.x * blockDim.x )
        {
            if ( x < nCaps )
            {
                const float4 sc = b.sectors[ x ];
                caps[ x ]       = make_float4( sc.x, sc.y, sc.z, cosf( sc.w ) );
            }
            if ( x >= nSeg )
                continue;
            const uint4 seg = b.segments[ x ];
            f3          ctr, nrm;
            float       rad;
            circle_world( b, seg.x, seg.y, ctr, nrm, rad );
            const f3 A = xyz( b.concavePos[ seg.z ] ) - ctr, B = xyz( b.concavePos[ seg.w ] ) - ctr;
            const f3 P = cross( nrm, A ), Q = cross( B, nrm );
            segGeo[ 4 * x ]     = make_float4( ctr.x, ctr.y, ctr.z, rad );
            segGeo[ 4 * x + 1 ] = make_float4( nrm.x, nrm.y, nrm.z, angle_ccw( A, B, nrm ) > 3.14159265f ? 1.f : 0.f );
            segGeo[ 4 * x + 2 ] = make_float4( P.x, P.y, P.z, 0.f );
            segGeo[ 4 * x + 3 ] = make_float4( Q.x, Q.y, Q.z, 0.f );
        }
    }

    // Caps of each atom, largest first (smallest cosine; a covered direction is then found early).
    // One wave per atom: up to 64 caps are ranked in registers (key, then index: a total order, so
    // ranks are distinct) and written to their places; longer lists are insertion-sorted by lane 0.
    // Caps of non-existent circles (contained spheres) carry a NaN cosine and are ranked last.
    __global__ void __launch_bounds__( 256 ) sdf_sort_caps_kernel( Buffers b, float4 * caps, const uint32_t * surfIds, const uint32_t * surfTotal )
    {
        // only faces that may be exposed are ever tested against caps: the surface atoms flagged
        // so in the compacted list, a wave per atom over a grid that strides the list
        const int      lane   = lane_id();
        const uint32_t total  = *surfTotal;
        const uint32_t stride = gridDim.x * ( blockDim.x / WaveSize );
        for ( uint32_t k = xcd_block() * ( blockDim.x / WaveSize ) + uint32_t( wave_in_block() ); k < total; k += stride )
        {
            const uint32_t id = surfIds[ k ];
            if ( !( id & SurfOpen ) )
                continue;
            const uint32_t a  = id & SurfAtom;
            const uint32_t c0 = b.vStart[ a ], c1 = min( b.vStart[ a + 1 ], b.capSectors );
            const int      n  = int( c1 > c0 ? c1 - c0 : 0u );
            auto           key = []( float w ) { return w == w ? w : 3.0f; };
            if ( n <= WaveSize )
            {
                const bool   live = lane < n;
                const float4 mine = live ? caps[ c0 + lane ] : make_float4( 0.f, 0.f, 0.f, 0.f );
                const float  kk   = key( mine.w );
                int          rank = 0;
                for ( int m = 0; m < n; m++ )
                {
                    const float km = __shfl( kk, m, WaveSize );
                    rank += ( km < kk || ( km == kk && m < lane ) ) ? 1 : 0;
                }
                wave_sync(); // every lane has read its cap before any is overwritten
                if ( live )
                    caps[ c0 + rank ] = mine;
                wave_sync();
                continue;
            }
            if ( lane == 0 )
                for ( uint32_t i = c0 + 1; i < c1; i++ )
                {
                    const float4 x = caps[ i ];
                    uint32_t     j = i;
                    for ( ; j > c0 && key( caps[ j - 1 ].w ) > key( x.w ); j-- )
                        caps[ j ] = caps[ j - 1 ];
                    caps[ j ] = x;
                }
        }
    }

    // Scattered points: a wave whose points fall in a compact block of cells is evaluated
    // cooperatively, any other wave lane by lane. Occupancy: the per-lane path sets the registers
    // (complete 79 VGPRs, six waves per SIMD; exterior 115, four); neither spills, and higher
    // bounds that spill ran no faster.
    template<bool EXT>
    __global__ void __launch_bounds__( 256, EXT ? 4 : 6 ) sdf_points_kernel( Buffers b, SdfArgs s )
    {
        __shared__ float4   stage[ 256 / WaveSize ][ StageSize ];
        __shared__ uint64_t masks[ 256 / WaveSize ][ 2 * MaskRows ];
        const uint64_t stride = uint64_t( gridDim.x ) * blockDim.x;
        for ( uint64_t base = uint64_t( blockIdx.x ) * blockDim.x; base < s.n; base += stride )
        {
            const uint64_t q = base + threadIdx.x;
            SdfLane        L;
            L.active = q < s.n;
            L.q      = L.active ? xyz( s.points[ q ] ) : mk( 0.f, 0.f, 0.f );
            int gx = 0, gy = 0, gz = 0;
            const Molecule M = cload( b.mols, 0 );
            if ( L.active )
                cell_of( M, L.q.x, L.q.y, L.q.z, gx, gy, gz );
            constexpr int Big = 1 << 30;
            const int     ex  = wave_max_int( L.active ? gx : -Big ) - wave_min_int( L.active ? gx : Big );
            const int     ey  = wave_max_int( L.active ? gy : -Big ) - wave_min_int( L.active ? gy : Big );
            const int     ez  = wave_max_int( L.active ? gz : -Big ) - wave_min_int( L.active ? gz : Big );
            uint64_t      code;
            float         v = 0.f;
            if ( !any_lane<true>( L.active ) )
                continue;
            if ( ex <= 2 && ey <= 2 && ez <= 2 )
                v = sdf_eval_lanes<true, EXT>( b, s, 0, L, code, stage[ wave_in_block() ], masks[ wave_in_block() ] );
            else if ( L.active )
                v = sdf_eval_lanes<false, EXT>( b, s, 0, L, code, nullptr, nullptr );
            if ( L.active )
            {
                const uint64_t dst = s.perm ? uint64_t( s.perm[ q ] ) : q;
                s.out[ dst ]       = v;
                if ( s.codes )
                    s.codes[ dst ] = code;
            }
        }
    }

    // Few points in caller order (under the sort threshold, scattered as the caller gave them): one
    // wave per point, evaluated cooperatively (the lanes share the point's stencil and test records
    // 64 at a time), so a small query spreads over the CUs instead of one wave walking 64 stencils
    // lane by lane. Lane 0 carries the point; the others only help.
    template<bool EXT>
    __global__ void __launch_bounds__( 256, EXT ? 5 : 8 ) sdf_point_waves_kernel( Buffers b, SdfArgs s )
    {
        __shared__ float4   stage[ 256 / WaveSize ][ StageSize ];
        __shared__ uint64_t masks[ 256 / WaveSize ][ 2 * MaskRows ];
        const uint64_t      q = uint64_t( blockIdx.x ) * ( blockDim.x / WaveSize ) + uint32_t( wave_in_block() );
        if ( q >= s.n )
            return;
        SdfLane L;
        L.active = lane_id() == 0;
        L.q      = xyz( s.points[ q ] );
        uint64_t    code;
        const float v = sdf_eval_lanes<true, EXT>( b, s, 0, L, code, stage[ wave_in_block() ], masks[ wave_in_block() ] );
        if ( L.active )
        {
            const uint64_t dst = s.perm ? uint64_t( s.perm[ q ] ) : q;
            s.out[ dst ]       = v;
            if ( s.codes )
                s.codes[ dst ] = code;
        }
    }

    // Cell key (molecule 0's brick) and index of every point, for the sort into cell order.
    __global__ void __launch_bounds__( 256 ) sdf_point_keys_kernel( Buffers b, const float4 * pts, uint32_t n, uint32_t * keys, uint32_t * idx )
    {
        const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
        if ( q >= n )
            return;
        const Molecule M = b.mols[ 0 ];
        const float4   p = pts[ q ];
        int            gx, gy, gz;
        cell_of( M, p.x, p.y, p.z, gx, gy, gz );
        keys[ q ] = cell_index( M, gx, gy, gz );
        idx[ q ]  = q;
    }

    __global__ void __launch_bounds__( 256 ) sdf_point_gather_kernel( const float4 * pts, const uint32_t * idx, uint32_t n, float4 * sorted )
    {
        const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
        if ( k < n )
            sorted[ k ] = pts[ idx[ k ] ];
    }

    // Regular grids: one wave per 4 x 4 x 4 brick of grid vertices (batches: per molecule grid).
    // Latency-bound, so occupancy decides: the complete field runs eight waves per SIMD (51
    // VGPRs), the exterior field five (94); neither spills. Uniform values (brick, molecule,
    // stencil, cell ranges, the chunk's records) live in SGPRs to keep VGPRs down.
    template<bool EXT>
    __global__ void __launch_bounds__( 256, EXT ? 5 : 8 ) sdf_grid_kernel( Buffers b, SdfArgs s, uint32_t molNb )
    {
        __shared__ float4   stage[ 256 / WaveSize ][ StageSize ];
        __shared__ uint64_t masks[ 256 / WaveSize ][ 2 * MaskRows ];
        // bricks of 4 x 4 x 4 vertices, one per wave (the launch has a wave for every brick); the
        // host refuses grids of more than 2^31 bricks (their output alone would exceed device
        // memory), so brick indices stay within 32 bits
        const uint32_t nbx = ( s.dx + 3 ) / 4, nby = ( s.dy + 3 ) / 4, nbz = ( s.dz + 3 ) / 4;
        const uint32_t bricks = nbx * nby * nbz, total = bricks * molNb;
        const uint64_t per    = uint64_t( s.dx ) * s.dy * s.dz;
        const int      lane   = lane_id();
        // the wave index is uniform (readfirstlane), so the brick, the molecule and its stencil
        // are scalar values
        const uint32_t w = blockIdx.x * ( blockDim.x / WaveSize ) + uint32_t( wave_in_block() );
        if ( w < total )
        {
            const uint32_t mol = w / bricks, br = w - mol * bricks;
            const uint32_t bxy = br % ( nbx * nby ), bz = br / ( nbx * nby );
            const uint32_t ix = ( bxy % nbx ) * 4 + ( lane & 3 ), iy = ( bxy / nbx ) * 4 + ( ( lane >> 2 ) & 3 ), iz = bz * 4 + ( lane >> 4 );
            SdfLane        L;
            L.active = ix < s.dx && iy < s.dy && iz < s.dz;
            const uint64_t q = uint64_t( mol ) * per + ( uint64_t( iz ) * s.dy + iy ) * s.dx + ix;
            float          ox = s.ox, oy = s.oy, oz = s.oz, sx = s.sx, sy = s.sy, sz = s.sz;
            if ( s.molGrids )
            {
                const float4 o = cload( s.molGrids, 2 * mol ), sp = cload( s.molGrids, 2 * mol + 1 );
                ox = o.x, oy = o.y, oz = o.z, sx = sp.x, sy = sp.y, sz = sp.z;
            }
            // lanes past the grid's edge stand at the molecule's first vertex
            L.q = L.active ? mk( ox + sx * float( ix ), oy + sy * float( iy ), oz + sz * float( iz ) ) : mk( ox, oy, oz );
            uint64_t    code;
            const float v = sdf_eval_lanes<true, EXT>( b, s, mol, L, code, stage[ wave_in_block() ], masks[ wave_in_block() ] );
            if ( L.active )
            {
                s.out[ q ] = v;
                if ( s.codes )
                    s.codes[ q ] = code;
            }
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

    // Gradient accumulation in fixed point: each contribution w * d is rounded to a multiple of
    // 2^-k, with 2^k = 2^30 / (max |w| rounded up to a power of two), and added as a 64-bit integer.
    // Integer sums do not depend on the order of the additions, so the gradient is the same in
    // every run; the resolution is 2^-30 of max |w| (finer than FP32 for gradients of order
    // max |w|), and sums stay below 2^63 up to 2^32 max |w| per atom.
    __device__ __forceinline__ double grad_scale( const unsigned long long * gf, uint32_t n )
    {
        const float maxW = __uint_as_float( uint32_t( gf[ 4 * size_t( n ) ] ) );
        int         e;
        (void)frexpf( maxW, &e ); // maxW < 2^e
        return ldexp( 1.0, 30 - e );
    }
    __device__ __forceinline__ void grad_add( unsigned long long * g, uint32_t a, d3 dc, double dr, float w, double scale )
    {
        const double v[ 4 ] = { double( w ) * dc.x, double( w ) * dc.y, double( w ) * dc.z, double( w ) * dr };
        for ( int c = 0; c < 4; c++ )
            atomicAdd( g + 4 * size_t( a ) + c, (unsigned long long)( long long )rint( v[ c ] * scale ) );
    }

    // max |w| over the queries (bits of a non-negative float order as unsigned integers)
    __global__ void grad_maxw_kernel( const float * w, uint64_t n, unsigned long long * gf, uint32_t atoms )
    {
        float m = 0.f;
        for ( uint64_t q = uint64_t( blockIdx.x ) * blockDim.x + threadIdx.x; q < n; q += uint64_t( gridDim.x ) * blockDim.x )
            m = fmaxf( m, fabsf( w[ q ] ) );
        for ( int o = 32; o > 0; o >>= 1 )
            m = fmaxf( m, __shfl_xor( m, o, WaveSize ) );
        if ( lane_id() == 0 && m > 0.f )
            atomicMax( gf + 4 * size_t( atoms ), (unsigned long long)__float_as_uint( m ) );
    }

    __global__ void grad_clear_kernel( unsigned long long * gf, uint32_t atoms )
    {
        for ( size_t a = size_t( blockIdx.x ) * blockDim.x + threadIdx.x; a < 4 * size_t( atoms ) + 2; a += size_t( gridDim.x ) * blockDim.x )
            gf[ a ] = 0ull;
    }

    // Gradient contributions of one query in FP32: the winner's nearest point p, e = (p - q) / |p - q|
    // and nu from the normal equations of its 1..3 sphere constraints (Cramer). False, with nothing
    // added, where FP32 is not enough: constraint vectors nearly dependent (the Gram determinant
    // below 1e-3 of the product of its diagonal: squared sine or squared volume) or, for an arc, q
    // near the circle's axis (the nearest point then turns fast with q). FP32 errors stay below 1e-4
    // of the contribution scale elsewhere.
    __device__ bool winner_gradient_fp32( const Buffers & b, const SdfArgs & s, f3 q, uint32_t type, uint32_t idx, float w, double scale )
    {
        const float rp = b.probeRadius;
        uint32_t    at[ 3 ];
        int         m = 0;
        f3          p;
        if ( type == SdfFace )
        {
            const float4 A = b.atoms[ idx ];
            const f3     v = q - xyz( A );
            at[ 0 ] = idx, m = 1;
            p       = xyz( A ) + v * ( ( A.w + rp ) / sqrtf( len2( v ) ) );
        }
        else if ( type == SdfVertex )
        {
            const int4 ids = b.concaveIds[ idx ];
            at[ 0 ] = uint32_t( ids.x ), at[ 1 ] = uint32_t( ids.y ), at[ 2 ] = uint32_t( ids.z ), m = 3;
            p = xyz( b.concavePos[ idx ] );
        }
        else
        {
            const uint2  ij = type == SdfCircle ? b.fullCircles[ idx ] : make_uint2( b.segments[ idx ].x, b.segments[ idx ].y );
            const float4 Ai = b.atoms[ ij.x ], Aj = b.atoms[ ij.y ];
            at[ 0 ] = ij.x, at[ 1 ] = ij.y, m = 2;
            const float Ri = Ai.w + rp, Rj = Aj.w + rp;
            const f3    dj  = xyz( Aj ) - xyz( Ai );
            const float d2  = len2( dj );
            const f3    nrm = dj * ( 1.f / sqrtf( d2 ) );
            const f3    ctr = xyz( Ai ) + dj * ( ( Ri * Ri - Rj * Rj + d2 ) / ( 2.f * d2 ) );
            const float rad = sqrtf( fmaxf( Ri * Ri - len2( ctr - xyz( Ai ) ), 0.f ) );
            const f3    v   = q - ctr;
            const f3    wv  = v - nrm * dot( v, nrm );
            const float lw  = sqrtf( len2( wv ) );
            if ( !( lw > 1e-2f * rad ) )
                return false;
            p = ctr + wv * ( rad / lw );
        }
        f3          e = p - q;
        const float D = sqrtf( len2( e ) );
        if ( !( D > 0.f ) )
            return false;
        e = e * ( 1.f / D );
        f3    a[ 3 ];
        float G[ 3 ][ 3 ], r[ 3 ], nu[ 3 ] = { 0.f, 0.f, 0.f };
        for ( int u = 0; u < m; u++ )
        {
            a[ u ] = p - xyz( b.atoms[ at[ u ] ] );
            r[ u ] = dot( a[ u ], e );
            for ( int v = 0; v <= u; v++ )
                G[ u ][ v ] = G[ v ][ u ] = dot( a[ u ], a[ v ] );
        }
        if ( m == 1 )
            nu[ 0 ] = r[ 0 ] / G[ 0 ][ 0 ];
        else if ( m == 2 )
        {
            const float det = G[ 0 ][ 0 ] * G[ 1 ][ 1 ] - G[ 0 ][ 1 ] * G[ 1 ][ 0 ];
            if ( !( det > 1e-3f * G[ 0 ][ 0 ] * G[ 1 ][ 1 ] ) )
                return false;
            nu[ 0 ] = ( r[ 0 ] * G[ 1 ][ 1 ] - G[ 0 ][ 1 ] * r[ 1 ] ) / det;
            nu[ 1 ] = ( G[ 0 ][ 0 ] * r[ 1 ] - G[ 1 ][ 0 ] * r[ 0 ] ) / det;
        }
        else
        {
            auto det3 = []( const float A[ 3 ][ 3 ] ) {
                return A[ 0 ][ 0 ] * ( A[ 1 ][ 1 ] * A[ 2 ][ 2 ] - A[ 1 ][ 2 ] * A[ 2 ][ 1 ] ) - A[ 0 ][ 1 ] * ( A[ 1 ][ 0 ] * A[ 2 ][ 2 ] - A[ 1 ][ 2 ] * A[ 2 ][ 0 ] ) +
                       A[ 0 ][ 2 ] * ( A[ 1 ][ 0 ] * A[ 2 ][ 1 ] - A[ 1 ][ 1 ] * A[ 2 ][ 0 ] );
            };
            const float d = det3( G );
            if ( !( d > 1e-3f * G[ 0 ][ 0 ] * G[ 1 ][ 1 ] * G[ 2 ][ 2 ] ) )
                return false;
            for ( int c = 0; c < 3; c++ )
            {
                float A[ 3 ][ 3 ];
                for ( int u = 0; u < 3; u++ )
                    for ( int v = 0; v < 3; v++ )
                        A[ u ][ v ] = v == c ? r[ u ] : G[ u ][ v ];
                nu[ c ] = det3( A ) / d;
            }
        }
        for ( int u = 0; u < m; u++ )
        {
            const float Ru = b.atoms[ at[ u ] ].w + rp;
            grad_add( s.gradFixed, at[ u ], mkd( a[ u ] * ( -nu[ u ] ) ), double( -nu[ u ] * Ru ), w, scale );
        }
        return true;
    }

    template<bool EXT>
    __global__ void __launch_bounds__( 256 ) sdf_backward_kernel( Buffers b, SdfArgs s )
    {
        const double rp    = b.probeRadius;
        const double scale = grad_scale( s.gradFixed, b.n );
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
                (void)sdf_eval<EXT>( b, s, qf, mol, code );
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
                grad_add( s.gradFixed, idx, v * ( -1.0 / dn ), -1.0, w, scale );
                continue;
            }
            if ( winner_gradient_fp32( b, s, qf, type, idx, w, scale ) )
                continue; // well conditioned: done in FP32
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
                grad_add( s.gradFixed, at[ u ], av[ u ] * ( -nu[ u ] ), -nu[ u ] * R( at[ u ] ), w, scale );
        }
    }

    // fixed point -> float, in sorted order (gradSorted) and in input order (gradInput)
    __global__ void gather_grad_kernel( const unsigned long long * gf, const uint32_t * s2i, uint32_t n, float4 * gradSorted, float4 * gradInput )
    {
        const uint32_t a = blockIdx.x * blockDim.x + threadIdx.x;
        if ( a >= n )
            return;
        const double inv = 1.0 / grad_scale( gf, n );
        float        v[ 4 ];
        for ( int c = 0; c < 4; c++ )
            v[ c ] = float( double( ( long long )gf[ 4 * size_t( a ) + c ] ) * inv );
        const float4 g = make_float4( v[ 0 ], v[ 1 ], v[ 2 ], v[ 3 ] );
        gradSorted[ a ]      = g;
        gradInput[ s2i[ a ] ] = g;
    }

} // namespace sesdf

using namespace sesdf;

namespace
{
    // Packed atoms and the compacted surface atoms of the current build (once per build).
    // With mayAlloc false, SESDF_ERR_CAPACITY instead of allocating missing or too small buffers.
    sesdf_status prepare_surface( Handle & h, bool mayAlloc )
    {
        SdfState &     d = h.sdf;
        const uint32_t n = h.n;
        if ( d.packAtoms && d.packSerial == h.buildSerial )
            return SESDF_OK;
        if ( !mayAlloc && ( !d.packAtoms || !d.geo || d.geoSectors < h.cap.sectors || d.geoSegments < h.cap.segments ) )
            return SESDF_ERR_CAPACITY;
        if ( !d.packAtoms )
        {
            const size_t n1    = size_t( n ) + 1;
            const size_t bytes = n1 * ( 16 + 16 + 32 + 8 ) + ( n1 * 4 ) * 4 + ( size_t( h.cellNb ) + 1 ) * 4 + 256;
            void *       p     = nullptr;
            if ( hipMalloc( &p, bytes ) != hipSuccess )
                return SESDF_ERR_HIP;
            h.sdfAllocs.push_back( p );
            h.sdfBytes += bytes;
            char * c        = static_cast<char *>( p );
            d.packAtoms     = reinterpret_cast<float4 *>( c ), c += n1 * 16;
            d.surfAtoms     = reinterpret_cast<float4 *>( c ), c += n1 * 16;
            d.surfRanges    = reinterpret_cast<int4 *>( c ), c += n1 * 32;
            d.surfCaps      = reinterpret_cast<int2 *>( c ), c += n1 * 8;
            d.surfIds       = reinterpret_cast<uint32_t *>( c ), c += n1 * 4;
            d.surfFlag      = reinterpret_cast<uint32_t *>( c ), c += n1 * 4;
            d.surfInfo      = reinterpret_cast<uint32_t *>( c ), c += n1 * 4;
            d.surfPrefix    = reinterpret_cast<uint32_t *>( c ), c += n1 * 4;
            d.surfCellStart = reinterpret_cast<uint32_t *>( c ), c += ( size_t( h.cellNb ) + 1 ) * 4;
            d.surfTotal     = reinterpret_cast<uint32_t *>( c );
        }
        // caps and segment geometry, sized by the build's capacities (which grow on overflow)
        if ( !d.geo || d.geoSectors < h.cap.sectors || d.geoSegments < h.cap.segments )
        {
            if ( d.geo )
            {
                h.sdfAllocs.erase( std::find( h.sdfAllocs.begin(), h.sdfAllocs.end(), static_cast<void *>( d.geo ) ) );
                (void)hipFree( d.geo );
                d.geo = nullptr;
            }
            const size_t bytes = ( size_t( h.cap.sectors ) + 4 * size_t( h.cap.segments ) ) * 16;
            void *       p     = nullptr;
            if ( hipMalloc( &p, bytes ) != hipSuccess )
                return SESDF_ERR_HIP;
            h.sdfAllocs.push_back( p );
            h.sdfBytes += bytes;
            d.geo         = static_cast<float4 *>( p );
            d.geoSectors  = h.cap.sectors;
            d.geoSegments = h.cap.segments;
        }
        hipStream_t s = h.stream;
        // sized by the capacities: the kernel reads the counts of the build from the device
        const uint32_t geoItems = std::max( h.cap.sectors, h.cap.segments );
        if ( geoItems )
            sdf_geometry_kernel<<<std::min<uint32_t>( ( geoItems + 255 ) / 256, uint32_t( device_cu_count() ) * 8 ), 256, 0, s>>>(
                h.b, d.geo, d.geo + d.geoSectors );
        if ( n )
            sdf_pack_kernel<<<( n + 3 ) / 4, 256, 0, s>>>( h.b, d.packAtoms, d.surfFlag, d.surfInfo );
        ScanArgs a {};
        a.in[ 0 ]     = d.surfFlag;
        a.out[ 0 ]    = d.surfPrefix;
        a.totals[ 0 ] = d.surfTotal;
        a.n           = n;
        a.ticket      = TicketSdf;
        a.stage       = 6;
        launch_scan<1>( h.b, a, s );
        const uint32_t items = std::max( n, h.cellNb + 1 );
        sdf_surface_kernel<<<( items + 255 ) / 256, 256, 0, s>>>( h.b, d.packAtoms, d.surfFlag, d.surfInfo, d.surfPrefix, d.surfAtoms,
                                                                 d.surfRanges, d.surfIds, d.surfCaps, d.surfCellStart );
        if ( n )
            sdf_sort_caps_kernel<<<std::min<uint32_t>( ( n + 3 ) / 4, uint32_t( device_cu_count() ) * 16 ), 256, 0, s>>>( h.b, d.geo, d.surfIds,
                                                                                                                      d.surfTotal );
        if ( hipGetLastError() != hipSuccess )
            return SESDF_ERR_HIP;
        d.packSerial = h.buildSerial;
        return SESDF_OK;
    }

    sesdf_status sdf_common( Handle & h, SdfArgs & s, int flags, bool mayAlloc = true )
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
        if ( h.n == 0 )
        {
            s.empty = 1; // no surface: nothing to prepare, every sample reads +band
            return SESDF_OK;
        }
        sesdf_status pst = prepare_surface( h, mayAlloc );
        if ( pst != SESDF_OK )
            return pst;
        s.sAtoms        = h.sdf.packAtoms;
        s.surfAtoms     = h.sdf.surfAtoms;
        s.surfRanges    = h.sdf.surfRanges;
        s.surfIds       = h.sdf.surfIds;
        s.surfCellStart = h.sdf.surfCellStart;
        s.surfCaps      = h.sdf.surfCaps;
        s.caps          = h.sdf.geo;
        s.segGeo        = h.sdf.geo + h.sdf.geoSectors;
        if ( s.exterior )
        {
            sesdf_status st = prepare_components( h, mayAlloc );
            if ( st != SESDF_OK )
                return st;
            s.comp = h.sdf.comp;
        }
        return SESDF_OK;
    }

    // +band (code: clamped) for every sample of an empty surface.
    __global__ void sdf_fill_empty_kernel( float * out, uint64_t * codes, uint64_t n, float band )
    {
        for ( uint64_t q = uint64_t( blockIdx.x ) * blockDim.x + threadIdx.x; q < n; q += uint64_t( gridDim.x ) * blockDim.x )
        {
            out[ q ] = band;
            if ( codes )
                codes[ q ] = uint64_t( SdfClamped ) << 40;
        }
    }

    // One wave per brick: brick costs differ by two orders of magnitude (outside vs surface), so
    // the hardware dispatcher, not a static stride, balances them. Grids beyond 2^31 blocks stride.
    uint64_t grid_bricks( const SdfArgs & s, uint32_t molNb )
    {
        return uint64_t( ( s.dx + 3 ) / 4 ) * ( ( s.dy + 3 ) / 4 ) * ( ( s.dz + 3 ) / 4 ) * molNb;
    }
    uint32_t brick_blocks( const SdfArgs & s, uint32_t molNb ) { return uint32_t( std::max<uint64_t>( 1, ( grid_bricks( s, molNb ) + 3 ) / 4 ) ); }
    // The grid kernel counts bricks in 32 bits: at most 2^31 (the output of 2^31 bricks, 549 GB,
    // already exceeds device memory).
    constexpr uint64_t MaxGridBricks = 1ull << 31;

    // One thread per item (costs vary per query): dynamic dispatch balances; beyond 2^31 blocks the
    // kernels stride.
    uint32_t grid_blocks( uint64_t n ) { return uint32_t( std::max<uint64_t>( 1, std::min<uint64_t>( ( n + 255 ) / 256, 0x7fffffffull ) ) ); }

    // Scattered points are evaluated in cell order: a wave of 64 consecutive points then falls in a
    // compact block of cells and runs cooperatively instead of lane by lane (a stable radix sort of
    // the cell keys; the results go back to the callers' order through the permutation).
    constexpr uint64_t PointSortMin = 8192;
    constexpr uint64_t PointSortMax = 0x7fffffffull; // the sort counts items in int

    // Buffers of the point sort: the points in cell order, two key and two index buffers of cap
    // words each, and the sort's histogram workspace.
    struct PointSortSpace
    {
        float4 *   sorted;
        uint32_t * words;
        uint32_t * hist;
        uint64_t   cap;
    };

    constexpr size_t point_sort_bytes( uint64_t cap ) { return cap * ( sizeof( float4 ) + 4 * sizeof( uint32_t ) ) + sort_hist_words( cap ) * 4; }

    PointSortSpace point_sort_space( void * p, uint64_t cap )
    {
        PointSortSpace w;
        w.sorted = static_cast<float4 *>( p );
        w.words  = reinterpret_cast<uint32_t *>( w.sorted + cap );
        w.hist   = w.words + 4 * cap;
        w.cap    = cap;
        return w;
    }

    void sort_points_in( Handle & h, SdfArgs & s, const PointSortSpace & w )
    {
        const uint32_t n  = uint32_t( s.n );
        hipStream_t    st = h.stream;
        int            bits = 1;
        while ( bits < 32 && ( uint64_t( 1 ) << bits ) < uint64_t( h.cellNb ) )
            bits++;
        uint32_t *k0 = w.words, *k1 = w.words + w.cap, *v0 = w.words + 2 * w.cap, *v1 = w.words + 3 * w.cap;
        sdf_point_keys_kernel<<<( n + 255 ) / 256, 256, 0, st>>>( h.b, s.points, n, k0, v0 );
        const int where = radix_sort_pairs( k0, v0, k1, v1, w.hist, n, bits, st );
        const uint32_t * perm = where ? v1 : v0;
        sdf_point_gather_kernel<<<( n + 255 ) / 256, 256, 0, st>>>( s.points, perm, n, w.sorted );
        s.points = w.sorted;
        s.perm   = perm;
    }

    sesdf_status sort_points( Handle & h, SdfArgs & s )
    {
        SdfState & d = h.sdf;
        if ( s.n > d.ptCap )
        {
            // grown on demand, kept with the other sampling buffers
            const uint64_t cap   = std::min<uint64_t>( std::max<uint64_t>( s.n, 2 * uint64_t( d.ptCap ) ), PointSortMax );
            const size_t   bytes = point_sort_bytes( cap );
            void *         p     = nullptr;
            if ( hipMalloc( &p, bytes ) != hipSuccess )
                return SESDF_ERR_CAPACITY;
            if ( d.ptSorted )
            {
                h.sdfAllocs.erase( std::find( h.sdfAllocs.begin(), h.sdfAllocs.end(), static_cast<void *>( d.ptSorted ) ) );
                h.sdfBytes -= point_sort_bytes( d.ptCap );
                (void)hipFree( d.ptSorted );
            }
            h.sdfAllocs.push_back( p );
            h.sdfBytes += bytes;
            d.ptSorted = static_cast<float4 *>( p );
            d.ptCap    = uint32_t( cap );
        }
        sort_points_in( h, s, point_sort_space( d.ptSorted, d.ptCap ) );
        return SESDF_OK;
    }
} // namespace

// Device array of the allocation-free sampling entry point: the point sort space for up to
// SESDF_NOALLOC_MAX_POINTS points. Each GPU holds its own copy; host code reaches it only through
// the address hipGetSymbolAddress returns on that GPU (the name is a host-side stand-in).
static __device__ __align__( 256 ) unsigned char scratch[ SESDF_NOALLOC_SCRATCH_BYTES ];

namespace
{
    static_assert( point_sort_bytes( SESDF_NOALLOC_MAX_POINTS ) == SESDF_NOALLOC_SCRATCH_BYTES, "scratch size" );

    std::mutex           g_scratchMutex;
    std::vector<void *>  g_scratchAddr; // per device ordinal

    // Address of `scratch` on the current device (the caller made h.device current), looked up on
    // the first call per device and kept.
    sesdf_status scratch_address( int device, void ** out )
    {
        std::lock_guard<std::mutex> lock( g_scratchMutex );
        if ( device < 0 )
            return SESDF_ERR_INVALID;
        if ( size_t( device ) >= g_scratchAddr.size() )
            g_scratchAddr.resize( size_t( device ) + 1, nullptr );
        if ( !g_scratchAddr[ device ] )
        {
            void *           p = nullptr;
            const hipError_t e = hipGetSymbolAddress( &p, HIP_SYMBOL( scratch ) );
            if ( e != hipSuccess || !p )
                return report_error( SESDF_ERR_HIP, "hipGetSymbolAddress(scratch) failed" );
            g_scratchAddr[ device ] = p;
        }
        *out = g_scratchAddr[ device ];
        return SESDF_OK;
    }
} // namespace

extern "C"
{
    sesdf_status sesdf_sample_points_noalloc( sesdf_handle * hh, const float * d_points, uint64_t npoints, float band, int flags,
                                              float * d_out, uint64_t * d_codes )
    {
        if ( !hh || ( npoints && ( !d_points || !d_out ) ) )
            return SESDF_ERR_INVALID;
        if ( npoints > SESDF_NOALLOC_MAX_POINTS )
            return report_error( SESDF_ERR_INVALID, "sesdf_sample_points_noalloc: npoints exceeds SESDF_NOALLOC_MAX_POINTS (1048576)" );
        Handle & h = *reinterpret_cast<Handle *>( hh );
        SdfArgs  s {};
        s.points        = reinterpret_cast<const float4 *>( d_points );
        s.n             = npoints;
        s.band          = band;
        s.out           = d_out;
        s.codes         = d_codes;
        sesdf_status st = sdf_common( h, s, flags, false );
        if ( st == SESDF_ERR_CAPACITY )
            return report_error( st, "sesdf_sample_points_noalloc: the sampling state of this build is not in place "
                                     "(call sesdf_sample_points_ex once with the same flags after the build)" );
        if ( st != SESDF_OK || npoints == 0 )
            return st;
        if ( s.empty )
            sdf_fill_empty_kernel<<<grid_blocks( npoints ), 256, 0, h.stream>>>( d_out, d_codes, npoints, band );
        else
        {
            if ( npoints >= PointSortMin )
            {
                void * p = nullptr;
                st       = scratch_address( h.device, &p );
                if ( st != SESDF_OK )
                    return st;
                sort_points_in( h, s, point_sort_space( p, SESDF_NOALLOC_MAX_POINTS ) );
            }
            if ( npoints < PointSortMin ) // a wave per point
            {
                const uint32_t blocks = uint32_t( ( npoints + 3 ) / 4 );
                if ( s.exterior )
                    sdf_point_waves_kernel<true><<<blocks, 256, 0, h.stream>>>( h.b, s );
                else
                    sdf_point_waves_kernel<false><<<blocks, 256, 0, h.stream>>>( h.b, s );
            }
            else if ( s.exterior )
                sdf_points_kernel<true><<<grid_blocks( npoints ), 256, 0, h.stream>>>( h.b, s );
            else
                sdf_points_kernel<false><<<grid_blocks( npoints ), 256, 0, h.stream>>>( h.b, s );
        }
        return hipGetLastError() == hipSuccess ? SESDF_OK : SESDF_ERR_HIP;
    }

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
        if ( s.empty )
            sdf_fill_empty_kernel<<<grid_blocks( npoints ), 256, 0, h.stream>>>( d_out, d_codes, npoints, band );
        else
        {
            if ( npoints >= PointSortMin && npoints <= PointSortMax )
            {
                sesdf_status ps = sort_points( h, s );
                if ( ps != SESDF_OK )
                    return ps;
            }
            if ( npoints < PointSortMin ) // a wave per point
            {
                const uint32_t blocks = uint32_t( ( npoints + 3 ) / 4 );
                if ( s.exterior )
                    sdf_point_waves_kernel<true><<<blocks, 256, 0, h.stream>>>( h.b, s );
                else
                    sdf_point_waves_kernel<false><<<blocks, 256, 0, h.stream>>>( h.b, s );
            }
            else if ( s.exterior )
                sdf_points_kernel<true><<<grid_blocks( npoints ), 256, 0, h.stream>>>( h.b, s );
            else
                sdf_points_kernel<false><<<grid_blocks( npoints ), 256, 0, h.stream>>>( h.b, s );
        }
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
        if ( s.empty )
        {
            *exterior_label = 0;
            return SESDF_OK;
        }
        if ( hipMemcpyAsync( host_labels, h.sdf.comp.atomLabel, size_t( h.n ) * 4, hipMemcpyDeviceToHost, h.stream ) != hipSuccess ||
             hipMemcpyAsync( exterior_label, h.sdf.comp.extComp, 4, hipMemcpyDeviceToHost, h.stream ) != hipSuccess ||
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
        if ( grid_bricks( s, 1 ) > MaxGridBricks )
            return report_error( SESDF_ERR_INVALID, "sesdf_sample_grid: more than 2^31 bricks of 4^3 vertices" );
        sesdf_status st = sdf_common( h, s, flags );
        if ( st != SESDF_OK || s.n == 0 )
            return st;
        if ( s.empty )
            sdf_fill_empty_kernel<<<grid_blocks( s.n ), 256, 0, h.stream>>>( d_out, nullptr, s.n, band );
        else if ( s.exterior )
            sdf_grid_kernel<true><<<brick_blocks( s, 1 ), 256, 0, h.stream>>>( h.b, s, 1u );
        else
            sdf_grid_kernel<false><<<brick_blocks( s, 1 ), 256, 0, h.stream>>>( h.b, s, 1u );
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
        if ( grid_bricks( s, uint32_t( M ) ) > MaxGridBricks )
            return report_error( SESDF_ERR_INVALID, "sesdf_batch_sample_grids: more than 2^31 bricks of 4^3 vertices" );
        sesdf_status st = sdf_common( h, s, 0 );
        if ( st != SESDF_OK )
            return st;
        if ( s.empty )
            sdf_fill_empty_kernel<<<grid_blocks( s.n ), 256, 0, h.stream>>>( d_out, nullptr, s.n, band );
        else if ( s.exterior )
            sdf_grid_kernel<true><<<brick_blocks( s, uint32_t( M ) ), 256, 0, h.stream>>>( h.b, s, uint32_t( M ) );
        else
            sdf_grid_kernel<false><<<brick_blocks( s, uint32_t( M ) ), 256, 0, h.stream>>>( h.b, s, uint32_t( M ) );
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
        if ( st != SESDF_OK || s.empty )
            return st; // no atoms: no gradient
        if ( s.recompute && npoints && npoints <= PointSortMax )
        {
            // winners re-derived by a cooperative forward pass into scratch codes (in cell order
            // from PointSortMin points, one wave per point below), instead of one thread per query:
            // the backward then reads them as stored ones
            SdfState & d = h.sdf;
            if ( npoints > d.ptEvalCap )
            {
                void * p = nullptr;
                if ( hipMalloc( &p, npoints * ( sizeof( uint64_t ) + sizeof( float ) ) ) != hipSuccess )
                    return SESDF_ERR_CAPACITY;
                h.sdfAllocs.push_back( p );
                h.sdfBytes += npoints * ( sizeof( uint64_t ) + sizeof( float ) );
                d.ptCodes   = static_cast<uint64_t *>( p );
                d.ptVals    = reinterpret_cast<float *>( d.ptCodes + npoints );
                d.ptEvalCap = npoints;
            }
            SdfArgs f   = s;
            f.out       = d.ptVals;
            f.codes     = d.ptCodes;
            f.weights   = nullptr;
            f.recompute = 0;
            if ( npoints >= PointSortMin )
            {
                sesdf_status ps = sort_points( h, f );
                if ( ps != SESDF_OK )
                    return ps;
                if ( f.exterior )
                    sdf_points_kernel<true><<<grid_blocks( npoints ), 256, 0, h.stream>>>( h.b, f );
                else
                    sdf_points_kernel<false><<<grid_blocks( npoints ), 256, 0, h.stream>>>( h.b, f );
            }
            else
            {
                const uint32_t blocks = uint32_t( ( npoints + 3 ) / 4 );
                if ( f.exterior )
                    sdf_point_waves_kernel<true><<<blocks, 256, 0, h.stream>>>( h.b, f );
                else
                    sdf_point_waves_kernel<false><<<blocks, 256, 0, h.stream>>>( h.b, f );
            }
            s.codes     = d.ptCodes;
            s.recompute = 0;
        }
        // fixed-point accumulators, kept with the handle (cleared by a kernel, not a memset node:
        // memset nodes of captured graphs are avoided)
        SdfState & d = h.sdf;
        if ( !d.gradFixed || d.gradFixedCap < h.n )
        {
            if ( d.gradFixed )
            {
                h.sdfAllocs.erase( std::find( h.sdfAllocs.begin(), h.sdfAllocs.end(), static_cast<void *>( d.gradFixed ) ) );
                (void)hipFree( d.gradFixed );
                d.gradFixed = nullptr;
            }
            void *       p     = nullptr;
            const size_t bytes = ( 4 * size_t( h.n ) + 2 ) * sizeof( unsigned long long );
            if ( hipMalloc( &p, bytes ) != hipSuccess )
                return SESDF_ERR_CAPACITY;
            h.sdfAllocs.push_back( p );
            h.sdfBytes += bytes;
            d.gradFixed    = static_cast<unsigned long long *>( p );
            d.gradFixedCap = h.n;
        }
        s.gradFixed = d.gradFixed;
        const uint32_t wide = uint32_t( device_cu_count() ) * 8;
        grad_clear_kernel<<<std::min<uint32_t>( uint32_t( ( 4 * size_t( h.n ) + 2 + 255 ) / 256 ), wide ), 256, 0, h.stream>>>( s.gradFixed, h.n );
        if ( npoints )
        {
            grad_maxw_kernel<<<std::min<uint32_t>( grid_blocks( npoints ), wide ), 256, 0, h.stream>>>( d_weights, npoints, s.gradFixed, h.n );
            if ( s.exterior )
                sdf_backward_kernel<true><<<grid_blocks( npoints ), 256, 0, h.stream>>>( h.b, s );
            else
                sdf_backward_kernel<false><<<grid_blocks( npoints ), 256, 0, h.stream>>>( h.b, s );
        }
        if ( h.n )
            gather_grad_kernel<<<( h.n + 255 ) / 256, 256, 0, h.stream>>>( s.gradFixed, h.b.sortedToInitial, h.n, s.gradSorted,
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
        if ( !h.sdf.gradScratch )
        {
            void * p = nullptr;
            if ( hipMalloc( &p, std::max<size_t>( size_t( h.n ) * 16, 16 ) ) != hipSuccess )
                return SESDF_ERR_HIP;
            h.sdfAllocs.push_back( p );
            h.sdfBytes += size_t( h.n ) * 16;
            h.sdf.gradScratch = static_cast<float4 *>( p );
        }
        return sesdf_sdf_backward_ex( hh, d_points, d_weights, npoints, band, 0, nullptr, reinterpret_cast<float *>( h.sdf.gradScratch ), d_grad );
    }
}
