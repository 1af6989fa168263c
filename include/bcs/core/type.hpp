// Value types of the host interface: spans, the atom vector, the bounding box, handle spans,
// ranges and a non-owning view of a device array. Same names, members and meaning as the
// reference's Span / ConstSpan / HandleSpan / Range (sesdf/include/bcs/core/type.hpp:9-53),
// Vec4f / Aabb (sesdf/include/bcs/core/math.hpp) and DeviceBuffer's accessors
// (sesdf/include/bcs/cuda/memory.cuh:119-160: get(), get<Type>(byteOffset), size<Type>(),
// toHost<Type>()), written for this library without glm or CUDA.
#ifndef BCS_CORE_TYPE_HPP
#define BCS_CORE_TYPE_HPP

#include <algorithm>
#include <cassert>
#include <cstddef>
#include <cstdint>
#include <vector>

#include "sesdf_c.h"

namespace bcs
{
    // xyz + vdW radius, 16 bytes (the layout of glm::vec4 that the reference reads).
    struct Vec4f
    {
        float x = 0.f, y = 0.f, z = 0.f, w = 0.f;
    };
    static_assert( sizeof( Vec4f ) == 16, "atoms are read as float4" );

    struct Vec3f
    {
        float x = 0.f, y = 0.f, z = 0.f;
    };

    template<class Type>
    struct Span
    {
        Type *      ptr  = nullptr;
        std::size_t size = 0;

        Span() = default;
        Span( Type * p, std::size_t n ) : ptr( p ), size( n ) {}
        Span( std::vector<Type> & v ) : ptr( v.data() ), size( v.size() ) {}

        Type * data() const { return ptr; }
        Type * begin() const { return ptr; }
        Type * end() const { return ptr + size; }
        Type & operator[]( std::size_t i ) const { return ptr[ i ]; }
    };

    template<class Type>
    struct ConstSpan
    {
        const Type * ptr  = nullptr;
        std::size_t  size = 0;

        ConstSpan() = default;
        ConstSpan( const Type * p, std::size_t n ) : ptr( p ), size( n ) {}
        ConstSpan( const std::vector<Type> & v ) : ptr( v.data() ), size( v.size() ) {}
        ConstSpan( const Span<Type> s ) : ptr( s.ptr ), size( s.size ) {}

        const Type * data() const { return ptr; }
        const Type * begin() const { return ptr; }
        const Type * end() const { return ptr + size; }
        const Type & operator[]( std::size_t i ) const { return ptr[ i ]; }
    };

    // Handle of a graphics buffer with an offset and a size in bytes (the *Graphics structs).
    template<class HandleType>
    struct HandleSpan
    {
        HandleType  handle {};
        std::size_t offset = 0;
        std::size_t size   = 0;

        HandleSpan() = default;
        HandleSpan( HandleType h, std::size_t o, std::size_t s ) : handle( h ), offset( o ), size( s ) {}
        HandleSpan( HandleType h, std::size_t s ) : handle( h ), offset( 0 ), size( s ) {}
    };

    template<class Type>
    struct RangeBase
    {
        Type start {};
        Type end {};

        Type size() const { return end - start; }
    };
    using Range = RangeBase<std::size_t>;

    struct Aabb
    {
        Vec3f min;
        Vec3f max;
    };

    // Box of the vdW spheres (centre -/+ radius), as bcs::getAabb (sesdf/src/core/molecule.cpp:132-143).
    inline Aabb getAabb( ConstSpan<Vec4f> molecule )
    {
        Aabb box;
        if ( molecule.size == 0 )
            return box;
        box.min = { 3.4e38f, 3.4e38f, 3.4e38f };
        box.max = { -3.4e38f, -3.4e38f, -3.4e38f };
        for ( const Vec4f & a : molecule )
        {
            box.min.x = std::min( box.min.x, a.x - a.w ), box.max.x = std::max( box.max.x, a.x + a.w );
            box.min.y = std::min( box.min.y, a.y - a.w ), box.max.y = std::max( box.max.y, a.y + a.w );
            box.min.z = std::min( box.min.z, a.z - a.w ), box.max.z = std::max( box.max.z, a.z + a.w );
        }
        return box;
    }

    // Non-owning view of a device array held by a builder (valid until its next build). The
    // accessors are the reference DeviceBuffer's: byte offsets in get<Type>, element counts in
    // size<Type>, and a blocking copy of the whole array in toHost<Type>.
    class DeviceBuffer
    {
      public:
        DeviceBuffer() = default;
        DeviceBuffer( void * ptr, std::size_t bytes ) : m_ptr( static_cast<uint8_t *>( ptr ) ), m_size( bytes ) {}

        uint8_t *       get() { return m_ptr; }
        const uint8_t * get() const { return m_ptr; }
        template<class Type>
        Type * get( std::size_t byteOffset = 0 )
        {
            return reinterpret_cast<Type *>( m_ptr + byteOffset );
        }
        template<class Type>
        const Type * get( std::size_t byteOffset = 0 ) const
        {
            return reinterpret_cast<const Type *>( m_ptr + byteOffset );
        }

        std::size_t size() const { return m_size; }
        template<class Type>
        std::size_t size() const
        {
            return m_size / sizeof( Type );
        }

        template<class Type>
        std::vector<Type> toHost() const
        {
            assert( m_size % sizeof( Type ) == 0 && "the array is not a whole number of these elements" );
            std::vector<Type> out( m_size / sizeof( Type ) );
            if ( !out.empty() )
                (void)sesdf_device_to_host( out.data(), m_ptr, out.size() * sizeof( Type ) );
            return out;
        }

        explicit operator bool() const { return m_ptr != nullptr; }

      private:
        uint8_t *   m_ptr  = nullptr;
        std::size_t m_size = 0;
    };
} // namespace bcs

#endif // BCS_CORE_TYPE_HPP
