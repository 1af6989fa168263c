// Value types of the host interface: a non-owning span, the atom vector, the bounding box and a
// non-owning device-array view. Same roles as the reference's ConstSpan / Vec4f / Aabb /
// DeviceBuffer (sesdf/include/bcs/core/type.hpp:21-36, sesdf/include/bcs/core/math.hpp:18-22,
// sesdf/include/bcs/cuda/memory.cuh), written for this library without glm or CUDA.
#ifndef BCS_CORE_TYPE_HPP
#define BCS_CORE_TYPE_HPP

#include <algorithm>
#include <cstddef>
#include <cstdint>
#include <vector>

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
    struct ConstSpan
    {
        const Type * ptr  = nullptr;
        std::size_t  size = 0;

        ConstSpan() = default;
        ConstSpan( const Type * p, std::size_t n ) : ptr( p ), size( n ) {}
        ConstSpan( const std::vector<Type> & v ) : ptr( v.data() ), size( v.size() ) {}

        const Type * data() const { return ptr; }
        const Type * begin() const { return ptr; }
        const Type * end() const { return ptr + size; }
        const Type & operator[]( std::size_t i ) const { return ptr[ i ]; }
    };

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

    // Non-owning view of a device array held by a builder (valid until its next build).
    class DeviceBuffer
    {
      public:
        DeviceBuffer() = default;
        DeviceBuffer( void * ptr, std::size_t bytes ) : m_ptr( ptr ), m_size( bytes ) {}

        template<class Type>
        Type * get() const
        {
            return static_cast<Type *>( m_ptr );
        }
        std::size_t size() const { return m_size; }

      private:
        void *      m_ptr  = nullptr;
        std::size_t m_size = 0;
    };

    // OpenGL buffer handle span of the reference's *Graphics structs (sesdf/include/bcs/core/type.hpp).
    template<class Handle>
    struct HandleSpan
    {
        Handle      id     = 0;
        std::size_t offset = 0;
        std::size_t size   = 0;
    };
} // namespace bcs

#endif // BCS_CORE_TYPE_HPP
