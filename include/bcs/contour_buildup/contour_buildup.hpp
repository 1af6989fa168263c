// bcs::ContourBuildup: the reference's Contour-Buildup surface method on an MI355X, over the C ABI
// of libsesdf_hip.so. Same constructor, build() and getGraphics() as the reference class
// (sesdf/include/bcs/contour_buildup/contour_buildup.hpp:24-43); header-only, link with -lsesdf_hip.
//
// Differences from the reference, by design:
//   - no OpenGL: getGraphics() reports the counts with vertex-array and texture names 0, and
//     getData() returns the device buffers those vertex arrays would bind;
//   - the device and stream are explicit (ContourBuildupOptions) instead of the max-GFLOPS device.
// Errors print one line to stderr and execution continues, as the reference's CUDA checks
// (sesdf/include/bcs/cuda/setup.inl:7-19); lastStatus() keeps the status of the last call.
#ifndef BCS_CONTOURBUILDUP_CONTOURBUILDUP_HPP
#define BCS_CONTOURBUILDUP_CONTOURBUILDUP_HPP

#include <cstdio>
#include <utility>
#include <vector>

#include "bcs/contour_buildup/graphics.hpp"
#include "bcs/core/type.hpp"
#include "sesdf_c.h"

namespace bcs
{
    struct ContourBuildupOptions
    {
        int      device   = 0;
        void *   stream   = nullptr; // hipStream_t; null: the object owns one
        uint32_t texSize  = 0;       // row length of the singularity texture; 0: 32768
        bool     useGraph = true;    // replay the stage sequence as one hipGraph
    };

    class ContourBuildup
    {
      public:
        ContourBuildup( const ConstSpan<Vec4f> molecule, const Aabb & aabb, const float probeRadius = 1.4f, uint32_t gridDimensions = 16,
                        bool buildSurface = true, const ContourBuildupOptions & options = ContourBuildupOptions {} )
            : m_molecule( molecule )
        {
            sesdf_cb_create_info info {};
            info.xyzr            = reinterpret_cast<const float *>( molecule.data() );
            info.atom_nb         = uint32_t( molecule.size );
            info.probe_radius    = probeRadius;
            info.aabb[ 0 ]       = aabb.min.x;
            info.aabb[ 1 ]       = aabb.min.y;
            info.aabb[ 2 ]       = aabb.min.z;
            info.aabb[ 3 ]       = aabb.max.x;
            info.aabb[ 4 ]       = aabb.max.y;
            info.aabb[ 5 ]       = aabb.max.z;
            info.grid_dimensions = gridDimensions;
            info.tex_size        = options.texSize;
            info.device          = options.device;
            info.stream          = options.stream;
            info.use_graph       = options.useGraph ? 1 : 0;
            info.build_surface   = buildSurface ? 1 : 0;
            check( sesdf_cb_create( &info, &m_handle ), "create" );
        }

        ContourBuildup( const ContourBuildup & )             = delete;
        ContourBuildup & operator=( const ContourBuildup & ) = delete;
        ContourBuildup( ContourBuildup && other ) noexcept { swap( other ); }
        ContourBuildup & operator=( ContourBuildup && other ) noexcept
        {
            if ( this != &other )
            {
                release();
                swap( other );
            }
            return *this;
        }
        ~ContourBuildup() { release(); }

        // Re-reads the molecule span given at construction and rebuilds every output.
        void build()
        {
            check( sesdf_cb_set_atoms( m_handle, reinterpret_cast<const float *>( m_molecule.data() ) ), "set_atoms" );
            check( sesdf_cb_build( m_handle ), "build" );
        }

        cb::ContourBuildupGraphics getGraphics() const
        {
            const sesdf_cb_counts      c = getCounts();
            cb::ContourBuildupGraphics g;
            g.atomNb              = c.atom_nb;
            g.sphericalTriangleNb = c.probe_nb;
            g.torusNb             = c.torus_nb;
            return g;
        }

        cb::ContourBuildupData getData() const
        {
            sesdf_cb_data d {};
            check( sesdf_cb_get_data( m_handle, &d ), "get_data" );
            cb::ContourBuildupData out;
            out.atomNb                 = d.atom_nb;
            out.atoms                  = reinterpret_cast<float4 *>( d.atoms );
            out.sphericalTriangleNb    = d.probe_nb;
            out.probePositions         = reinterpret_cast<float4 *>( d.probe_positions );
            out.sphereTriangleVec1     = reinterpret_cast<float4 *>( d.sphere_triangle_vec1 );
            out.sphereTriangleVec2     = reinterpret_cast<float4 *>( d.sphere_triangle_vec2 );
            out.sphereTriangleVec3     = reinterpret_cast<float4 *>( d.sphere_triangle_vec3 );
            out.torusNb                = d.torus_nb;
            out.torusPositions         = reinterpret_cast<float4 *>( d.torus_positions );
            out.torusAxes              = reinterpret_cast<float4 *>( d.torus_axes );
            out.torusVisibilitySpheres = reinterpret_cast<float4 *>( d.torus_visibility_spheres );
            out.texCoords              = d.tex_coords;
            out.singularityTexture     = d.singularity_texture;
            out.probeNeighborNb        = d.probe_neighbor_nb;
            out.texSize                = d.tex_size;
            out.probeRadius            = d.probe_radius;
            return out;
        }

        // Counts of the last build (probes, tori, truncated neighbour lists, capacities, memory).
        sesdf_cb_counts getCounts() const
        {
            sesdf_cb_counts c {};
            check( sesdf_cb_get_counts( m_handle, &c ), "get_counts" );
            return c;
        }
        sesdf_status      lastStatus() const { return m_status; }
        sesdf_cb_handle * handle() const { return m_handle; }

      private:
        sesdf_status check( sesdf_status st, const char * what ) const
        {
            m_status = st;
            if ( st != SESDF_OK )
                std::fprintf( stderr, "bcs::ContourBuildup %s: %s (%s)\n", what, sesdf_status_string( st ), sesdf_last_error() );
            return st;
        }
        void swap( ContourBuildup & other ) noexcept
        {
            std::swap( m_molecule, other.m_molecule );
            std::swap( m_handle, other.m_handle );
            std::swap( m_status, other.m_status );
        }
        void release()
        {
            if ( m_handle )
                sesdf_cb_destroy( m_handle );
            m_handle = nullptr;
        }

        ConstSpan<Vec4f>      m_molecule {};
        sesdf_cb_handle *     m_handle = nullptr;
        mutable sesdf_status  m_status = SESDF_OK;
    };
} // namespace bcs

#endif // BCS_CONTOURBUILDUP_CONTOURBUILDUP_HPP
