// Molecule loading from MMTF files: the atoms of the first model, each with the van der Waals
// radius of its element. This is what the reference's bcs::load (sesdf/src/core/loader.cpp:14-71,
// chemfiles' MMTF reader) hands to the builders: every atom of every residue of the first frame in
// file order, the element upper-cased and looked up in the radius table of
// sesdf/src/core/molecule.cpp:9-131 (unknown elements: 1.20 A).
//
// The decoder is written for this library from the MMTF 1.0 specification (RCSB): the file is a
// MessagePack map; per-atom and per-group arrays are binary blobs with a 12-byte big-endian header
// (codec, element count, parameter) followed by the encoded data.
#include <algorithm>
#include <cctype>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "handle.h"
#include "sesdf_c.h"

namespace
{
    // -- radius table: element symbols by atomic number (the reference's SYMBOL names) ------------
    struct Element
    {
        const char * symbol;
        float        radius;
    };
    // Z = 1 .. 118; index 0 is the unknown element.
    constexpr Element Elements[] = {
        { "", 1.20f },    { "H", 1.20f },   { "HE", 1.43f },  { "LI", 2.12f },  { "BE", 1.98f },  { "B", 1.91f },   { "C", 1.77f },
        { "N", 1.66f },   { "O", 1.50f },   { "F", 1.46f },   { "NE", 1.58f },  { "NA", 2.50f },  { "MG", 2.51f },  { "AL", 2.25f },
        { "SI", 2.19f },  { "P", 1.90f },   { "S", 1.89f },   { "CL", 1.82f },  { "AR", 1.83f },  { "K", 2.73f },   { "CA", 2.62f },
        { "SC", 2.58f },  { "TI", 2.46f },  { "V", 2.42f },   { "CR", 2.45f },  { "MN", 2.45f },  { "FE", 2.44f },  { "CO", 2.40f },
        { "NI", 2.40f },  { "CU", 2.38f },  { "ZN", 2.39f },  { "GA", 2.32f },  { "GE", 2.29f },  { "AS", 1.88f },  { "SE", 1.82f },
        { "BR", 1.86f },  { "KR", 2.25f },  { "RB", 3.21f },  { "SR", 2.84f },  { "Y", 2.75f },   { "ZR", 2.52f },  { "NB", 2.56f },
        { "MO", 2.45f },  { "TC", 2.44f },  { "RU", 2.46f },  { "RH", 2.44f },  { "PD", 2.15f },  { "AG", 2.53f },  { "CD", 2.49f },
        { "IN", 2.43f },  { "SN", 2.42f },  { "SB", 2.47f },  { "TE", 1.99f },  { "I", 2.04f },   { "XE", 2.06f },  { "CS", 3.48f },
        { "BA", 3.03f },  { "LA", 2.98f },  { "CE", 2.88f },  { "PR", 2.92f },  { "ND", 2.95f },  { "PM", 0.00f },  { "SM", 2.90f },
        { "EU", 2.87f },  { "GD", 2.83f },  { "TB", 2.79f },  { "DY", 2.87f },  { "HO", 2.81f },  { "ER", 2.83f },  { "TM", 2.79f },
        { "YB", 2.80f },  { "LU", 2.74f },  { "HF", 2.63f },  { "TA", 2.53f },  { "W", 2.57f },   { "RE", 2.49f },  { "OS", 2.48f },
        { "IR", 2.41f },  { "PT", 2.29f },  { "AU", 2.32f },  { "HG", 2.45f },  { "TL", 2.47f },  { "PB", 2.60f },  { "BI", 2.54f },
        { "PO", 0.00f },  { "AT", 0.00f },  { "RN", 0.00f },  { "FR", 0.00f },  { "RA", 0.00f },  { "AC", 2.80f },  { "TH", 2.93f },
        { "PA", 2.88f },  { "U", 2.71f },   { "NP", 2.82f },  { "PU", 2.81f },  { "AM", 2.83f },  { "CM", 3.05f },  { "BK", 3.40f },
        { "CF", 3.05f },  { "ES", 2.70f },  { "FM", 0.00f },  { "MD", 0.00f },  { "NO", 0.00f },  { "LR", 0.00f },  { "RF", 0.00f },
        { "DD", 0.00f },  { "SG", 0.00f },  { "BHJ", 0.00f }, { "HS", 0.00f },  { "MT", 0.00f },  { "DS", 0.00f },  { "RG", 0.00f },
        { "UUB", 0.00f }, { "UUT", 0.00f }, { "UUQ", 0.00f }, { "UUP", 0.00f }, { "UUH", 0.00f }, { "UUS", 0.00f }, { "UUO", 0.00f },
    };
    static_assert( sizeof( Elements ) / sizeof( Elements[ 0 ] ) == 119, "unknown + 118 elements" );

    float element_radius( const std::string & element )
    {
        std::string up = element;
        std::transform( up.begin(), up.end(), up.begin(), []( unsigned char c ) { return char( std::toupper( c ) ); } );
        if ( up.empty() )
            return Elements[ 0 ].radius;
        for ( const Element & e : Elements )
            if ( up == e.symbol )
                return e.radius;
        return Elements[ 0 ].radius;
    }

    // -- MessagePack reading ----------------------------------------------------------------------
    struct Reader
    {
        const uint8_t * p;
        const uint8_t * end;

        void need( size_t n ) const
        {
            if ( size_t( end - p ) < n )
                throw std::runtime_error( "truncated MessagePack data" );
        }
        uint64_t be( int bytes )
        {
            need( size_t( bytes ) );
            uint64_t v = 0;
            for ( int i = 0; i < bytes; i++ )
                v = ( v << 8 ) | p[ i ];
            p += bytes;
            return v;
        }
        uint8_t tag()
        {
            need( 1 );
            return *p++;
        }

        // Integer of any MessagePack integer form.
        int64_t integer()
        {
            const uint8_t t = tag();
            if ( t <= 0x7f )
                return t;
            if ( t >= 0xe0 )
                return int8_t( t );
            switch ( t )
            {
            case 0xcc: return int64_t( be( 1 ) );
            case 0xcd: return int64_t( be( 2 ) );
            case 0xce: return int64_t( be( 4 ) );
            case 0xcf: return int64_t( be( 8 ) );
            case 0xd0: return int8_t( be( 1 ) );
            case 0xd1: return int16_t( be( 2 ) );
            case 0xd2: return int32_t( be( 4 ) );
            case 0xd3: return int64_t( be( 8 ) );
            default: throw std::runtime_error( "MessagePack integer expected" );
            }
        }

        // Length of a string or binary blob (MMTF writers use both for encoded arrays).
        size_t blob_length( uint8_t t )
        {
            if ( t >= 0xa0 && t <= 0xbf )
                return t & 0x1f;
            switch ( t )
            {
            case 0xc4:
            case 0xd9: return size_t( be( 1 ) );
            case 0xc5:
            case 0xda: return size_t( be( 2 ) );
            case 0xc6:
            case 0xdb: return size_t( be( 4 ) );
            default: throw std::runtime_error( "MessagePack string or binary expected" );
            }
        }
        std::string blob()
        {
            const size_t n = blob_length( tag() );
            need( n );
            std::string s( reinterpret_cast<const char *>( p ), n );
            p += n;
            return s;
        }

        size_t array_length()
        {
            const uint8_t t = tag();
            if ( t >= 0x90 && t <= 0x9f )
                return t & 0x0f;
            if ( t == 0xdc )
                return size_t( be( 2 ) );
            if ( t == 0xdd )
                return size_t( be( 4 ) );
            throw std::runtime_error( "MessagePack array expected" );
        }
        size_t map_length()
        {
            const uint8_t t = tag();
            if ( t >= 0x80 && t <= 0x8f )
                return t & 0x0f;
            if ( t == 0xde )
                return size_t( be( 2 ) );
            if ( t == 0xdf )
                return size_t( be( 4 ) );
            throw std::runtime_error( "MessagePack map expected" );
        }

        void skip()
        {
            const uint8_t t = tag();
            if ( t <= 0x7f || t >= 0xe0 || t == 0xc0 || t == 0xc2 || t == 0xc3 )
                return;
            if ( ( t >= 0x80 && t <= 0x8f ) || t == 0xde || t == 0xdf )
            {
                --p;
                const size_t n = map_length();
                for ( size_t i = 0; i < 2 * n; i++ )
                    skip();
                return;
            }
            if ( ( t >= 0x90 && t <= 0x9f ) || t == 0xdc || t == 0xdd )
            {
                --p;
                const size_t n = array_length();
                for ( size_t i = 0; i < n; i++ )
                    skip();
                return;
            }
            if ( ( t >= 0xa0 && t <= 0xbf ) || ( t >= 0xc4 && t <= 0xc6 ) || ( t >= 0xd9 && t <= 0xdb ) )
            {
                const size_t n = blob_length( t );
                need( n );
                p += n;
                return;
            }
            switch ( t )
            {
            case 0xca: need( 4 ), p += 4; return;
            case 0xcb: need( 8 ), p += 8; return;
            case 0xcc:
            case 0xd0: need( 1 ), p += 1; return;
            case 0xcd:
            case 0xd1: need( 2 ), p += 2; return;
            case 0xce:
            case 0xd2: need( 4 ), p += 4; return;
            case 0xcf:
            case 0xd3: need( 8 ), p += 8; return;
            case 0xd4: need( 2 ), p += 2; return; // fixext 1, 2, 4, 8, 16: type byte + data
            case 0xd5: need( 3 ), p += 3; return;
            case 0xd6: need( 5 ), p += 5; return;
            case 0xd7: need( 9 ), p += 9; return;
            case 0xd8: need( 17 ), p += 17; return;
            case 0xc7:
            case 0xc8:
            case 0xc9:
            {
                const size_t n = size_t( be( t == 0xc7 ? 1 : t == 0xc8 ? 2 : 4 ) );
                need( n + 1 );
                p += n + 1;
                return;
            }
            default: throw std::runtime_error( "unknown MessagePack tag" );
            }
        }

        std::vector<int64_t> int_array()
        {
            const size_t         n = array_length();
            std::vector<int64_t> v( n );
            for ( size_t i = 0; i < n; i++ )
                v[ i ] = integer();
            return v;
        }
    };

    // -- MMTF array codecs ------------------------------------------------------------------------
    struct Encoded
    {
        int32_t     codec = 0, count = 0, param = 0;
        std::string data;
    };

    Encoded parse_encoded( const std::string & blob )
    {
        if ( blob.size() < 12 )
            throw std::runtime_error( "MMTF encoded array shorter than its header" );
        auto be32 = [ & ]( size_t o ) {
            const uint8_t * b = reinterpret_cast<const uint8_t *>( blob.data() ) + o;
            return int32_t( ( uint32_t( b[ 0 ] ) << 24 ) | ( uint32_t( b[ 1 ] ) << 16 ) | ( uint32_t( b[ 2 ] ) << 8 ) | uint32_t( b[ 3 ] ) );
        };
        Encoded e;
        e.codec = be32( 0 ), e.count = be32( 4 ), e.param = be32( 8 );
        e.data  = blob.substr( 12 );
        return e;
    }

    // Big-endian signed integers of `bytes` bytes each.
    std::vector<int64_t> be_ints( const std::string & d, int bytes )
    {
        const size_t         n = d.size() / size_t( bytes );
        std::vector<int64_t> v( n );
        const uint8_t *      b = reinterpret_cast<const uint8_t *>( d.data() );
        for ( size_t i = 0; i < n; i++ )
        {
            uint32_t u = 0;
            for ( int k = 0; k < bytes; k++ )
                u = ( u << 8 ) | b[ i * bytes + k ];
            v[ i ] = bytes == 1 ? int64_t( int8_t( u ) ) : bytes == 2 ? int64_t( int16_t( u ) ) : int64_t( int32_t( u ) );
        }
        return v;
    }

    std::vector<int64_t> run_length( const std::vector<int64_t> & pairs )
    {
        std::vector<int64_t> out;
        for ( size_t i = 0; i + 1 < pairs.size(); i += 2 )
        {
            if ( pairs[ i + 1 ] < 0 )
                throw std::runtime_error( "negative run length" );
            out.insert( out.end(), size_t( pairs[ i + 1 ] ), pairs[ i ] );
        }
        return out;
    }

    std::vector<int64_t> delta( std::vector<int64_t> v )
    {
        for ( size_t i = 1; i < v.size(); i++ )
            v[ i ] = int64_t( int32_t( uint32_t( v[ i - 1 ] ) + uint32_t( v[ i ] ) ) );
        return v;
    }

    // Values at the extremes of the packed type continue into the next one.
    std::vector<int64_t> recursive_index( const std::vector<int64_t> & v, int64_t lo, int64_t hi )
    {
        std::vector<int64_t> out;
        int64_t              acc = 0;
        for ( int64_t x : v )
        {
            acc += x;
            if ( x != lo && x != hi )
            {
                out.push_back( acc );
                acc = 0;
            }
        }
        return out;
    }

    std::vector<int64_t> decode_ints( const Encoded & e )
    {
        switch ( e.codec )
        {
        case 2: return be_ints( e.data, 1 );
        case 3: return be_ints( e.data, 2 );
        case 4: return be_ints( e.data, 4 );
        case 7: return run_length( be_ints( e.data, 4 ) );
        case 8: return delta( run_length( be_ints( e.data, 4 ) ) );
        case 14: return recursive_index( be_ints( e.data, 2 ), -32768, 32767 );
        case 15: return recursive_index( be_ints( e.data, 1 ), -128, 127 );
        default: throw std::runtime_error( "MMTF integer codec " + std::to_string( e.codec ) + " is not supported" );
        }
    }

    // Integer-encoded floats: value / divisor, rounded once to single precision.
    std::vector<float> decode_floats( const Encoded & e )
    {
        std::vector<float> out;
        if ( e.codec == 1 )
        {
            for ( int64_t u : be_ints( e.data, 4 ) )
            {
                const uint32_t bits = uint32_t( u );
                float          f;
                std::memcpy( &f, &bits, 4 );
                out.push_back( f );
            }
            return out;
        }
        std::vector<int64_t> ints;
        switch ( e.codec )
        {
        case 9: ints = run_length( be_ints( e.data, 4 ) ); break;
        case 10: ints = delta( recursive_index( be_ints( e.data, 2 ), -32768, 32767 ) ); break;
        case 11: ints = be_ints( e.data, 2 ); break;
        case 12: ints = recursive_index( be_ints( e.data, 2 ), -32768, 32767 ); break;
        case 13: ints = recursive_index( be_ints( e.data, 1 ), -128, 127 ); break;
        default: throw std::runtime_error( "MMTF float codec " + std::to_string( e.codec ) + " is not supported" );
        }
        if ( e.param == 0 )
            throw std::runtime_error( "MMTF float codec with a zero divisor" );
        out.resize( ints.size() );
        for ( size_t i = 0; i < ints.size(); i++ )
            out[ i ] = float( ints[ i ] ) / float( e.param );
        return out;
    }

    std::vector<int64_t> int_field( Reader & r )
    {
        // an encoded blob, or (for the per-model and per-chain counts) a plain array
        const uint8_t t = *r.p;
        if ( ( t >= 0x90 && t <= 0x9f ) || t == 0xdc || t == 0xdd )
            return r.int_array();
        return decode_ints( parse_encoded( r.blob() ) );
    }

    std::vector<float> load_mmtf( const char * path )
    {
        FILE * f = std::fopen( path, "rb" );
        if ( !f )
            throw std::runtime_error( std::string( "cannot open " ) + path );
        std::vector<uint8_t> bytes;
        uint8_t              buf[ 1 << 16 ];
        size_t               got;
        while ( ( got = std::fread( buf, 1, sizeof( buf ), f ) ) > 0 )
            bytes.insert( bytes.end(), buf, buf + got );
        std::fclose( f );

        Reader                                r { bytes.data(), bytes.data() + bytes.size() };
        std::vector<int64_t>                  chainsPerModel, groupsPerChain, groupType;
        std::vector<std::vector<std::string>> groupElements;
        std::vector<float>                    coord[ 3 ];
        int64_t                               numAtoms = -1;
        const size_t                          fields   = r.map_length();
        for ( size_t k = 0; k < fields; k++ )
        {
            const std::string key = r.blob();
            if ( key == "numAtoms" )
                numAtoms = r.integer();
            else if ( key == "chainsPerModel" )
                chainsPerModel = int_field( r );
            else if ( key == "groupsPerChain" )
                groupsPerChain = int_field( r );
            else if ( key == "groupTypeList" )
                groupType = int_field( r );
            else if ( key == "xCoordList" || key == "yCoordList" || key == "zCoordList" )
                coord[ key[ 0 ] - 'x' ] = decode_floats( parse_encoded( r.blob() ) );
            else if ( key == "groupList" )
            {
                const size_t ng = r.array_length();
                groupElements.resize( ng );
                for ( size_t g = 0; g < ng; g++ )
                {
                    const size_t nf = r.map_length();
                    for ( size_t q = 0; q < nf; q++ )
                    {
                        const std::string gk = r.blob();
                        if ( gk == "elementList" )
                        {
                            const size_t na = r.array_length();
                            groupElements[ g ].resize( na );
                            for ( size_t a = 0; a < na; a++ )
                                groupElements[ g ][ a ] = r.blob();
                        }
                        else
                            r.skip();
                    }
                }
            }
            else
                r.skip();
        }
        if ( numAtoms < 0 || chainsPerModel.empty() || coord[ 0 ].size() != size_t( numAtoms ) || coord[ 1 ].size() != size_t( numAtoms ) ||
             coord[ 2 ].size() != size_t( numAtoms ) )
            throw std::runtime_error( "MMTF file lacks atoms, models or coordinates" );

        // the first model: its chains, their groups, the groups' atoms in order
        int64_t chains = chainsPerModel[ 0 ], groups = 0;
        if ( chains < 0 || size_t( chains ) > groupsPerChain.size() )
            throw std::runtime_error( "MMTF chain counts are inconsistent" );
        for ( int64_t c = 0; c < chains; c++ )
            groups += groupsPerChain[ size_t( c ) ];
        if ( groups < 0 || size_t( groups ) > groupType.size() )
            throw std::runtime_error( "MMTF group counts are inconsistent" );
        std::vector<float> xyzr;
        size_t             atom = 0;
        for ( int64_t g = 0; g < groups; g++ )
        {
            const int64_t t = groupType[ size_t( g ) ];
            if ( t < 0 || size_t( t ) >= groupElements.size() )
                throw std::runtime_error( "MMTF group type out of range" );
            for ( const std::string & el : groupElements[ size_t( t ) ] )
            {
                if ( atom >= size_t( numAtoms ) )
                    throw std::runtime_error( "MMTF groups hold more atoms than numAtoms" );
                xyzr.insert( xyzr.end(), { coord[ 0 ][ atom ], coord[ 1 ][ atom ], coord[ 2 ][ atom ], element_radius( el ) } );
                atom++;
            }
        }
        return xyzr;
    }
} // namespace

extern "C"
{
    sesdf_status sesdf_load_mmtf( const char * path, float ** xyzr, uint64_t * atom_nb )
    {
        if ( !path || !xyzr || !atom_nb )
            return SESDF_ERR_INVALID;
        *xyzr    = nullptr;
        *atom_nb = 0;
        try
        {
            std::vector<float> v = load_mmtf( path );
            float *            out = static_cast<float *>( std::malloc( std::max<size_t>( v.size(), 1 ) * sizeof( float ) ) );
            if ( !out )
                return sesdf::report_error( SESDF_ERR_CAPACITY, "sesdf_load_mmtf: out of host memory" );
            std::copy( v.begin(), v.end(), out );
            *xyzr    = out;
            *atom_nb = v.size() / 4;
            return SESDF_OK;
        }
        catch ( const std::exception & e )
        {
            return sesdf::report_error( SESDF_ERR_INVALID, ( std::string( "sesdf_load_mmtf: " ) + e.what() ).c_str() );
        }
    }

    void sesdf_free_atoms( float * xyzr ) { std::free( xyzr ); }

    float sesdf_element_radius( const char * element ) { return element_radius( element ? element : "" ); }
}
