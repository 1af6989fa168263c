// Synthetic atom clouds for tests and benchmarks (host C++, no GPU).
//
// Every workload is generated from a project-owned PRNG (xoshiro256** seeded
// through SplitMix64) so the same seed gives the same atoms on every machine.
// Atoms are written as float4 xyzr (array of structs, 16 B per atom), the layout
// the reference host classes take (bcs::ConstSpan<Vec4f>, sesdf/include/bcs/sesdf/sesdf.hpp:18).
//
// Construction of a "globule": rejection sampling uniform in a ball of radius
// R = (3N / (4 pi rho))^(1/3), rho = 0.05 atoms/A^3, with a minimum pair distance
// of 1.2 A; element mix C 62% / N 17% / O 19% / S 2% with van der Waals radii
// 1.70 / 1.55 / 1.52 / 1.80 A.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace
{
    inline uint64_t splitmix64( uint64_t & s )
    {
        uint64_t z = ( s += 0x9E3779B97F4A7C15ull );
        z          = ( z ^ ( z >> 30 ) ) * 0xBF58476D1CE4E5B9ull;
        z          = ( z ^ ( z >> 27 ) ) * 0x94D049BB133111EBull;
        return z ^ ( z >> 31 );
    }

    struct Rng
    {
        uint64_t s[ 4 ];
        explicit Rng( uint64_t seed )
        {
            uint64_t x = seed;
            for ( auto & v : s )
                v = splitmix64( x );
        }
        static inline uint64_t rotl( uint64_t x, int k ) { return ( x << k ) | ( x >> ( 64 - k ) ); }
        uint64_t               next()
        {
            const uint64_t result = rotl( s[ 1 ] * 5, 7 ) * 9;
            const uint64_t t      = s[ 1 ] << 17;
            s[ 2 ] ^= s[ 0 ];
            s[ 3 ] ^= s[ 1 ];
            s[ 1 ] ^= s[ 2 ];
            s[ 0 ] ^= s[ 3 ];
            s[ 2 ] ^= t;
            s[ 3 ] = rotl( s[ 3 ], 45 );
            return result;
        }
        // uniform double in [0, 1)
        double uniform() { return static_cast<double>( next() >> 11 ) * ( 1.0 / 9007199254740992.0 ); }
    };

    uint64_t subSeed( uint64_t seed, uint64_t stream )
    {
        uint64_t x = seed ^ ( 0xD1B54A32D192ED03ull * ( stream + 1 ) );
        return splitmix64( x );
    }

    constexpr double Density     = 0.05;
    constexpr double MinDistance = 1.2;

    double globuleRadius( uint32_t n ) { return std::cbrt( 3.0 * n / ( 4.0 * M_PI * Density ) ); }

    float drawRadius( Rng & rng )
    {
        const double u = rng.uniform();
        if ( u < 0.62 )
            return 1.70f; // C
        if ( u < 0.79 )
            return 1.55f; // N
        if ( u < 0.98 )
            return 1.52f; // O
        return 1.80f;     // S
    }

    struct Void
    {
        double x, y, z, r;
    };

    // Dense hash grid with MinDistance cells over [-R, R]^3, used for the minimum-distance rejection.
    struct RejectGrid
    {
        double              lo;
        int                 dim;
        std::vector<int>    head;
        std::vector<int>    next;
        std::vector<double> pts;

        RejectGrid( double R, uint32_t n ) : lo( -R - MinDistance ), dim( int( std::ceil( ( 2 * R + 2 * MinDistance ) / MinDistance ) ) + 1 )
        {
            head.assign( size_t( dim ) * dim * dim, -1 );
            next.reserve( n );
            pts.reserve( size_t( n ) * 3 );
        }
        int cell( double v ) const
        {
            int c = int( std::floor( ( v - lo ) / MinDistance ) );
            return c < 0 ? 0 : ( c >= dim ? dim - 1 : c );
        }
        bool free( double x, double y, double z ) const
        {
            const int cx = cell( x ), cy = cell( y ), cz = cell( z );
            for ( int dz = -1; dz <= 1; dz++ )
                for ( int dy = -1; dy <= 1; dy++ )
                    for ( int dx = -1; dx <= 1; dx++ )
                    {
                        const int ix = cx + dx, iy = cy + dy, iz = cz + dz;
                        if ( ix < 0 || iy < 0 || iz < 0 || ix >= dim || iy >= dim || iz >= dim )
                            continue;
                        for ( int p = head[ ( size_t( iz ) * dim + iy ) * dim + ix ]; p >= 0; p = next[ p ] )
                        {
                            const double ddx = pts[ 3 * p ] - x, ddy = pts[ 3 * p + 1 ] - y, ddz = pts[ 3 * p + 2 ] - z;
                            if ( ddx * ddx + ddy * ddy + ddz * ddz < MinDistance * MinDistance )
                                return false;
                        }
                    }
            return true;
        }
        void add( double x, double y, double z )
        {
            const int    id = int( next.size() );
            const size_t c  = ( size_t( cell( z ) ) * dim + cell( y ) ) * dim + cell( x );
            next.push_back( head[ c ] );
            head[ c ] = id;
            pts.push_back( x );
            pts.push_back( y );
            pts.push_back( z );
        }
    };

    // Globule centred on the origin; returns the number of atoms written.
    uint32_t makeGlobule( uint32_t n, Rng & rng, const std::vector<Void> & voids, double * xyz, float * radii )
    {
        const double R = globuleRadius( n );
        RejectGrid   grid( R, n );
        uint32_t     accepted = 0;
        uint64_t     attempts = 0;
        while ( accepted < n )
        {
            if ( ++attempts > 1000ull * n + 1000000ull )
                break; // cannot place more atoms (over-full construction)
            const double x = R * ( 2 * rng.uniform() - 1 );
            const double y = R * ( 2 * rng.uniform() - 1 );
            const double z = R * ( 2 * rng.uniform() - 1 );
            if ( x * x + y * y + z * z > R * R )
                continue;
            bool inVoid = false;
            for ( const Void & v : voids )
            {
                const double dx = x - v.x, dy = y - v.y, dz = z - v.z;
                if ( dx * dx + dy * dy + dz * dz < v.r * v.r )
                {
                    inVoid = true;
                    break;
                }
            }
            if ( inVoid || !grid.free( x, y, z ) )
                continue;
            grid.add( x, y, z );
            xyz[ 3 * accepted + 0 ] = x;
            xyz[ 3 * accepted + 1 ] = y;
            xyz[ 3 * accepted + 2 ] = z;
            radii[ accepted ]       = drawRadius( rng );
            accepted++;
        }
        return accepted;
    }
} // namespace

extern "C"
{
    // One globule of n atoms centred at the origin (float4 xyzr into out[4n]).
    // Returns the number of atoms placed (n unless the construction is over-full).
    uint32_t sesdf_synth_globule( uint32_t n, uint64_t seed, float * out )
    {
        Rng                 rng( seed );
        std::vector<double> xyz( size_t( n ) * 3 );
        std::vector<float>  r( n );
        const uint32_t      got = makeGlobule( n, rng, {}, xyz.data(), r.data() );
        for ( uint32_t i = 0; i < got; i++ )
        {
            out[ 4 * i + 0 ] = float( xyz[ 3 * i + 0 ] );
            out[ 4 * i + 1 ] = float( xyz[ 3 * i + 1 ] );
            out[ 4 * i + 2 ] = float( xyz[ 3 * i + 2 ] );
            out[ 4 * i + 3 ] = r[ i ];
        }
        return got;
    }

    // Globule of n atoms with nVoids empty spherical voids of radius U(4, 8) A, centres uniform in
    // the ball of radius R - Rv - 5 A and mutually non-overlapping; atoms inside a void are rejected
    // until exactly n are accepted. voidsOut[4 * nVoids] receives (cx, cy, cz, radius).
    uint32_t sesdf_synth_globule_voids( uint32_t n, uint32_t nVoids, uint64_t seed, float * out, float * voidsOut )
    {
        Rng               rng( seed );
        const double      R = globuleRadius( n );
        std::vector<Void> voids;
        uint32_t          guard = 0;
        while ( voids.size() < nVoids && guard++ < 100000u )
        {
            const double rv   = 4.0 + 4.0 * rng.uniform();
            const double span = R - rv - 5.0;
            const double x = span * ( 2 * rng.uniform() - 1 ), y = span * ( 2 * rng.uniform() - 1 ),
                         z = span * ( 2 * rng.uniform() - 1 );
            if ( x * x + y * y + z * z > span * span )
                continue;
            bool overlap = false;
            for ( const Void & v : voids )
            {
                const double dx = x - v.x, dy = y - v.y, dz = z - v.z;
                if ( std::sqrt( dx * dx + dy * dy + dz * dz ) < rv + v.r )
                    overlap = true;
            }
            if ( !overlap )
                voids.push_back( { x, y, z, rv } );
        }
        for ( size_t v = 0; v < voids.size(); v++ )
        {
            voidsOut[ 4 * v + 0 ] = float( voids[ v ].x );
            voidsOut[ 4 * v + 1 ] = float( voids[ v ].y );
            voidsOut[ 4 * v + 2 ] = float( voids[ v ].z );
            voidsOut[ 4 * v + 3 ] = float( voids[ v ].r );
        }
        std::vector<double> xyz( size_t( n ) * 3 );
        std::vector<float>  r( n );
        const uint32_t      got = makeGlobule( n, rng, voids, xyz.data(), r.data() );
        for ( uint32_t i = 0; i < got; i++ )
        {
            out[ 4 * i + 0 ] = float( xyz[ 3 * i + 0 ] );
            out[ 4 * i + 1 ] = float( xyz[ 3 * i + 1 ] );
            out[ 4 * i + 2 ] = float( xyz[ 3 * i + 2 ] );
            out[ 4 * i + 3 ] = r[ i ];
        }
        return got;
    }

    // nGlobules globules of atomsPer atoms each; centres on a Fibonacci sphere of radius shellRadius
    // (y = 1 - 2 (g + 1/2) / G, phi = g pi (3 - sqrt 5)), each rotated by a uniform random rotation
    // (Shoemake quaternion). out receives nGlobules * atomsPer float4 atoms, globule-major.
    uint32_t sesdf_synth_fibonacci( uint32_t nGlobules, uint32_t atomsPer, float shellRadius, uint64_t seed, float * out )
    {
        uint32_t            total = 0;
        std::vector<double> xyz( size_t( atomsPer ) * 3 );
        std::vector<float>  r( atomsPer );
        for ( uint32_t g = 0; g < nGlobules; g++ )
        {
            Rng            rng( subSeed( seed, g ) );
            const uint32_t got = makeGlobule( atomsPer, rng, {}, xyz.data(), r.data() );

            const double u1 = rng.uniform(), u2 = rng.uniform(), u3 = rng.uniform();
            const double qx = std::sqrt( 1 - u1 ) * std::sin( 2 * M_PI * u2 );
            const double qy = std::sqrt( 1 - u1 ) * std::cos( 2 * M_PI * u2 );
            const double qz = std::sqrt( u1 ) * std::sin( 2 * M_PI * u3 );
            const double qw = std::sqrt( u1 ) * std::cos( 2 * M_PI * u3 );
            // rotation matrix of the unit quaternion (qw, qx, qy, qz)
            const double m00 = 1 - 2 * ( qy * qy + qz * qz ), m01 = 2 * ( qx * qy - qz * qw ), m02 = 2 * ( qx * qz + qy * qw );
            const double m10 = 2 * ( qx * qy + qz * qw ), m11 = 1 - 2 * ( qx * qx + qz * qz ), m12 = 2 * ( qy * qz - qx * qw );
            const double m20 = 2 * ( qx * qz - qy * qw ), m21 = 2 * ( qy * qz + qx * qw ), m22 = 1 - 2 * ( qx * qx + qy * qy );

            const double y   = 1.0 - 2.0 * ( g + 0.5 ) / nGlobules;
            const double rr  = std::sqrt( std::max( 0.0, 1.0 - y * y ) );
            const double phi = g * M_PI * ( 3.0 - std::sqrt( 5.0 ) );
            const double cx = shellRadius * std::cos( phi ) * rr, cy = shellRadius * y, cz = shellRadius * std::sin( phi ) * rr;

            for ( uint32_t i = 0; i < got; i++ )
            {
                const double px = xyz[ 3 * i ], py = xyz[ 3 * i + 1 ], pz = xyz[ 3 * i + 2 ];
                float *      o  = out + 4 * size_t( total + i );
                o[ 0 ]          = float( cx + m00 * px + m01 * py + m02 * pz );
                o[ 1 ]          = float( cy + m10 * px + m11 * py + m12 * pz );
                o[ 2 ]          = float( cz + m20 * px + m21 * py + m22 * pz );
                o[ 3 ]          = r[ i ];
            }
            total += got;
        }
        return total;
    }

    // Atom counts of a batch: round(exp(U(ln lo, ln hi))) for each of nMolecules molecules.
    void sesdf_synth_batch_sizes( uint32_t nMolecules, uint32_t lo, uint32_t hi, uint64_t seed, uint32_t * sizes )
    {
        Rng          rng( seed );
        const double a = std::log( double( lo ) ), b = std::log( double( hi ) );
        for ( uint32_t m = 0; m < nMolecules; m++ )
            sizes[ m ] = uint32_t( std::llround( std::exp( a + ( b - a ) * rng.uniform() ) ) );
    }

    // Batch of globules: molecule m has sizes[m] atoms generated from sub-stream m of seed; all are
    // centred at the origin. out receives sum(sizes) float4 atoms, molecule-major.
    uint64_t sesdf_synth_batch( uint32_t nMolecules, const uint32_t * sizes, uint64_t seed, float * out )
    {
        uint64_t total = 0;
        for ( uint32_t m = 0; m < nMolecules; m++ )
            total += sesdf_synth_globule( sizes[ m ], subSeed( seed, m ), out + 4 * total );
        return total;
    }

    // Uniform doubles / normal variates for query-point and weight generation.
    void sesdf_synth_uniform( uint64_t n, uint64_t seed, double * out )
    {
        Rng rng( seed );
        for ( uint64_t i = 0; i < n; i++ )
            out[ i ] = rng.uniform();
    }
    void sesdf_synth_normal( uint64_t n, uint64_t seed, double * out )
    {
        Rng rng( seed );
        for ( uint64_t i = 0; i < n; i += 2 )
        {
            // Box-Muller
            double u1 = rng.uniform(), u2 = rng.uniform();
            if ( u1 < 1e-300 )
                u1 = 1e-300;
            const double m = std::sqrt( -2.0 * std::log( u1 ) );
            out[ i ]       = m * std::cos( 2 * M_PI * u2 );
            if ( i + 1 < n )
                out[ i + 1 ] = m * std::sin( 2 * M_PI * u2 );
        }
    }
}
