// bcs::Ssesdf: exterior solvent-excluded surface (torus records + concave patches) of one molecule
// on an MI355X, over the C ABI of libsesdf_hip.so. Same constructor, build() and getters as the
// reference class (sesdf/include/bcs/ssesdf/ssesdf.hpp:14-77); header-only, link with -lsesdf_hip.
// As for bcs::Sesdf, graphics = true creates no OpenGL buffers (getGraphics() describes the
// reference's buffer layout, packGraphics() fills device buffers in it) and the device and stream
// are explicit.
#ifndef BCS_SSESDF_SSESDF_HPP
#define BCS_SSESDF_SSESDF_HPP

#include <vector>

#include "bcs/core/builder.hpp"
#include "bcs/ssesdf/graphics.hpp"

namespace bcs
{
    class Ssesdf
    {
      public:
        Ssesdf( ConstSpan<Vec4f> molecule, const Aabb & aabb, const float probeRadius = 1.4f, bool buildSurface = true, bool graphics = true,
                const BuildOptions & options = BuildOptions {} )
            : m_builder( molecule, aabb, probeRadius, buildSurface, SESDF_MODE_EXTERIOR, options )
        {
            (void)graphics; // accepted for source compatibility: there is no GL interop
        }

        Ssesdf( const Ssesdf & )             = delete;
        Ssesdf & operator=( const Ssesdf & ) = delete;
        Ssesdf( Ssesdf && )                  = default;
        Ssesdf & operator=( Ssesdf && )      = default;
        ~Ssesdf()                            = default;

        void build() { m_builder.build(); }

        // Field layout of the reference's shared GL buffers (sesdf/src/ssesdf/ssesdf.cu:189-210): ids[0..2]
        // for the atom, circle and concave-patch buffers (0 without ids), byte offsets and sizes.
        ssesdf::SsesdfGraphics getGraphics( const GLuint * ids = nullptr ) const
        {
            sesdf_graphics_layout L {};
            m_builder.check( sesdf_get_graphics_layout( m_builder.handle(), &L ), "get_graphics_layout" );
            const sesdf_counts c = m_builder.counts();
            auto span = [ & ]( int q ) {
                const sesdf_graphics_span & sp = L.spans[ q ];
                return HandleSpan<GLuint>( ids ? ids[ sp.buffer ] : 0u, std::size_t( sp.offset ), std::size_t( sp.size ) );
            };
            ssesdf::SsesdfGraphics g;
            g.atomNb                  = c.atom_nb;
            g.atoms                   = span( 0 );
            g.circlePatchNb           = c.torus_nb;
            g.circlePosition          = span( 1 );
            g.circleAxis              = span( 2 );
            g.circleVs                = span( 3 );
            g.concavePatchNb          = c.concave_patch_nb;
            g.concavePatchesPosition  = span( 4 );
            g.concavePatchesId        = span( 5 );
            g.concavePatchesNeighbors = span( 6 );
            g.probeRadius             = m_builder.probeRadius();
            g.maxConcaveNeighbors     = uint16_t( m_builder.internal().max_intersection_neighbors );
            return g;
        }
        sesdf_status packGraphics( void * const * dBuffers ) const
        {
            return m_builder.check( sesdf_pack_graphics( m_builder.handle(), dBuffers ), "pack_graphics" );
        }

        ssesdf::SsesdfData getData()
        {
            ssesdf_data d {};
            m_builder.check( sesdf_get_exterior_data( m_builder.handle(), &d ), "get_exterior_data" );
            ssesdf::SsesdfData out;
            out.atomNb                  = d.atom_nb;
            out.atoms                   = reinterpret_cast<float4 *>( d.atoms );
            out.circlePatchNb           = d.circle_patch_nb;
            out.circlePosition          = reinterpret_cast<float4 *>( d.circle_position );
            out.circleAxis              = reinterpret_cast<float4 *>( d.circle_axis );
            out.circleVs                = reinterpret_cast<float4 *>( d.circle_vs );
            out.concavePatchNb          = d.concave_patch_nb;
            out.concavePatchesPosition  = reinterpret_cast<float4 *>( d.concave_patches_position );
            out.concavePatchesId        = reinterpret_cast<int4 *>( d.concave_patches_id );
            out.concavePatchesNeighbors = reinterpret_cast<float4 *>( d.concave_patches_neighbors );
            out.probeRadius             = d.probe_radius;
            out.maxConcaveNeighbors     = d.max_concave_neighbors;
            return out;
        }

        // Helpers to analyse the surface (sesdf/include/bcs/ssesdf/ssesdf.hpp:35-41).
        uint16_t              getMaxNeighborPerAtom() const { return uint16_t( m_builder.internal().max_neighbors ); }
        uint16_t              getMaxIntersectionsPerCircles() const { return m_builder.internal().max_intersections_per_circle; }
        uint16_t              getMaxIntersectionNeighbors() const { return m_builder.internal().max_intersection_neighbors; }
        std::vector<uint32_t> getAtomIndices() { return m_builder.atomIndices(); }
        DeviceBuffer &        getDAtomIndices() { return refresh().m_dAtomIndices; }
        DeviceBuffer &        getDAtomNeighborsCount() { return refresh().m_dNeighborsCount; }
        DeviceBuffer &        getDAtomNeighborsIndices() { return refresh().m_dNeighborsIndices; }
        DeviceBuffer &        getDCircleVisibilityStatus() { return refresh().m_dStatus; }

        sesdf_counts   getCounts() const { return m_builder.counts(); }
        sesdf_status   lastStatus() const { return m_builder.lastStatus(); }
        sesdf_handle * handle() const { return m_builder.handle(); }

      private:
        Ssesdf & refresh()
        {
            const sesdf_internal in = m_builder.internal();
            const std::size_t    n  = m_builder.counts().atom_nb;
            m_dAtomIndices          = DeviceBuffer( in.sorted_to_initial, n * 4 );
            m_dNeighborsCount       = DeviceBuffer( in.neighbor_count, n * 4 );
            m_dNeighborsIndices     = DeviceBuffer( in.neighbor_ids, n * in.max_neighbors * 4 );
            m_dStatus               = DeviceBuffer( in.circle_status, n * in.max_neighbors );
            return *this;
        }

        detail::Builder m_builder;
        DeviceBuffer    m_dAtomIndices, m_dNeighborsCount, m_dNeighborsIndices, m_dStatus;
    };
} // namespace bcs

#endif // BCS_SSESDF_SSESDF_HPP
