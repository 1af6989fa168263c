// Move-only owner of one libsesdf_hip handle, shared by bcs::Sesdf and bcs::Ssesdf.
//
// Error behaviour follows the reference's CUDA checks (sesdf/include/bcs/cuda/setup.inl:7-19):
// a failing call prints one line to stderr and execution continues; the status of the last call
// is also kept (lastStatus()) for callers that want to test it.
#ifndef BCS_CORE_BUILDER_HPP
#define BCS_CORE_BUILDER_HPP

#include <cstdio>
#include <utility>
#include <vector>

#include "bcs/core/type.hpp"
#include "sesdf_c.h"

namespace bcs
{
    // Launch environment and capacities the reference fixes at compile time
    // (BCS_SESDF_MAXNEIGHBORPERATOM, sesdf/include/bcs/sesdf/sesdf.hpp:8-10).
    struct BuildOptions
    {
        int      device                   = 0;
        void *   stream                   = nullptr; // hipStream_t; null: the builder owns one
        uint32_t maxNeighbors             = 128;     // per-atom neighbour capacity (multiple of 32, <= 320)
        uint32_t maxIntersectionNeighbors = 32;
        float    cellSize                 = 0.f;     // grid cell edge in A; 0: default
        bool     useGraph                 = true;    // replay the stage sequence as one hipGraph
    };

    namespace detail
    {
        class Builder
        {
          public:
            Builder() = default;
            Builder( ConstSpan<Vec4f> molecule, const Aabb & aabb, float probeRadius, bool buildSurface, int mode,
                     const BuildOptions & options )
                : m_molecule( molecule ), m_probeRadius( probeRadius )
            {
                sesdf_create_info info {};
                info.xyzr                       = reinterpret_cast<const float *>( molecule.data() );
                info.atom_nb                    = uint32_t( molecule.size );
                info.probe_radius               = probeRadius;
                info.aabb[ 0 ]                  = aabb.min.x;
                info.aabb[ 1 ]                  = aabb.min.y;
                info.aabb[ 2 ]                  = aabb.min.z;
                info.aabb[ 3 ]                  = aabb.max.x;
                info.aabb[ 4 ]                  = aabb.max.y;
                info.aabb[ 5 ]                  = aabb.max.z;
                info.mode                       = mode;
                info.max_neighbors              = options.maxNeighbors;
                info.max_intersection_neighbors = options.maxIntersectionNeighbors;
                info.cell_size                  = options.cellSize;
                info.device                     = options.device;
                info.stream                     = options.stream;
                info.graphics                   = 0; // no OpenGL interop: getGraphics() reports sizes only
                info.use_graph                  = options.useGraph ? 1 : 0;
                info.build_surface              = buildSurface ? 1 : 0;
                check( sesdf_create( &info, &m_handle ), "create" );
            }
            Builder( const Builder & )             = delete;
            Builder & operator=( const Builder & ) = delete;
            Builder( Builder && other ) noexcept { swap( other ); }
            Builder & operator=( Builder && other ) noexcept
            {
                if ( this != &other )
                {
                    release();
                    swap( other );
                }
                return *this;
            }
            ~Builder() { release(); }

            void build()
            {
                check( sesdf_set_atoms( m_handle, reinterpret_cast<const float *>( m_molecule.data() ) ), "set_atoms" );
                check( sesdf_build( m_handle ), "build" );
            }

            sesdf_counts counts() const
            {
                sesdf_counts c {};
                check( sesdf_get_counts( m_handle, &c ), "get_counts" );
                return c;
            }
            sesdf_internal internal() const
            {
                sesdf_internal in {};
                check( sesdf_get_internal( m_handle, &in ), "get_internal" );
                return in;
            }
            std::vector<uint32_t> atomIndices() const
            {
                std::vector<uint32_t> out( m_molecule.size );
                check( sesdf_get_atom_indices( m_handle, out.data() ), "get_atom_indices" );
                return out;
            }

            sesdf_handle * handle() const { return m_handle; }
            float          probeRadius() const { return m_probeRadius; }
            sesdf_status   lastStatus() const { return m_last; }

            sesdf_status check( sesdf_status s, const char * what ) const
            {
                m_last = s;
                if ( s != SESDF_OK )
                    std::fprintf( stderr, "bcs::sesdf %s: %s (%s)\n", what, sesdf_status_string( s ), sesdf_last_error() );
                return s;
            }

          private:
            void release()
            {
                if ( m_handle )
                    sesdf_destroy( m_handle );
                m_handle = nullptr;
            }
            void swap( Builder & o ) noexcept
            {
                std::swap( m_handle, o.m_handle );
                std::swap( m_molecule, o.m_molecule );
                std::swap( m_probeRadius, o.m_probeRadius );
                std::swap( m_last, o.m_last );
            }

            sesdf_handle *       m_handle = nullptr;
            ConstSpan<Vec4f>     m_molecule;
            float                m_probeRadius = 0.f;
            mutable sesdf_status m_last        = SESDF_OK;
        };
    } // namespace detail
} // namespace bcs

#endif // BCS_CORE_BUILDER_HPP
