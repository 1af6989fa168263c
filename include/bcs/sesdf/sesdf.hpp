// bcs::Sesdf: complete solvent-excluded surface of one molecule on an MI355X, over the C ABI of
// libsesdf_hip.so. Same constructor, build() and getters as the reference class
// (sesdf/include/bcs/sesdf/sesdf.hpp:15-88); header-only, link with -lsesdf_hip.
//
// Differences from the reference, by design:
//   - graphics = true does not create OpenGL buffers (no GL interop): getGraphics() returns the
//     reference's buffer layout (offsets, sizes; ids given by the caller or 0) and packGraphics()
//     fills caller-owned device buffers in that layout; getData() returns the device pointers;
//   - the device and stream are explicit (BuildOptions) instead of the max-GFLOPS device;
//   - SDF sampling and its gradient are added (sampleGrid / samplePoints / sdfBackward).
#ifndef BCS_SESDF_SESDF_HPP
#define BCS_SESDF_SESDF_HPP

#include <vector>

#include "bcs/core/builder.hpp"
#include "bcs/sesdf/graphics.hpp"

namespace bcs
{
    class Sesdf
    {
      public:
        Sesdf( ConstSpan<Vec4f> molecule, const Aabb & aabb, const float probeRadius = 1.4f, bool buildSurface = true, bool graphics = true,
               const BuildOptions & options = BuildOptions {} )
            : m_builder( molecule, aabb, probeRadius, buildSurface, SESDF_MODE_COMPLETE, options )
        {
            (void)graphics; // accepted for source compatibility: there is no GL interop
        }

        Sesdf( const Sesdf & )             = delete;
        Sesdf & operator=( const Sesdf & ) = delete;
        Sesdf( Sesdf && )                  = default;
        Sesdf & operator=( Sesdf && )      = default;
        ~Sesdf()                           = default;

        // Re-reads the molecule span given at construction and rebuilds every output.
        void build() { m_builder.build(); }

        // Field layout of the reference's shared GL buffers (sesdf/src/sesdf/sesdf.cu:266-299): one
        // handle per buffer (ids[0..4]: atoms, convex patches, full circles + sectors, segments,
        // concave patches; 0 without ids), byte offsets and sizes. packGraphics() fills device
        // buffers laid out so.
        sesdf::SesdfGraphics getGraphics( const GLuint * ids = nullptr ) const
        {
            sesdf_graphics_layout L {};
            m_builder.check( sesdf_get_graphics_layout( m_builder.handle(), &L ), "get_graphics_layout" );
            const sesdf_counts   c = m_builder.counts();
            auto span = [ & ]( int q ) {
                const sesdf_graphics_span & sp = L.spans[ q ];
                return HandleSpan<GLuint>( ids ? ids[ sp.buffer ] : 0u, std::size_t( sp.offset ), std::size_t( sp.size ) );
            };
            sesdf::SesdfGraphics g;
            g.atomNb                  = c.atom_nb;
            g.atoms                   = span( 0 );
            g.convexPatchNb           = c.atom_nb;
            g.convexPatches           = span( 1 );
            g.sectorNb                = c.sector_nb;
            g.sectors                 = span( 2 );
            g.segmentPatchNb          = c.segment_nb;
            g.segmentPatches          = span( 3 );
            g.circlePatchNb           = c.full_circle_nb;
            g.circlePatches           = span( 4 );
            g.concavePatchNb          = c.concave_patch_nb;
            g.concavePatchesPosition  = span( 5 );
            g.concavePatchesId        = span( 6 );
            g.concavePatchesNeighbors = span( 7 );
            g.probeRadius             = m_builder.probeRadius();
            g.maxConcaveNeighbors     = uint16_t( m_builder.internal().max_intersection_neighbors );
            return g;
        }
        sesdf_graphics_layout getGraphicsLayout() const
        {
            sesdf_graphics_layout L {};
            m_builder.check( sesdf_get_graphics_layout( m_builder.handle(), &L ), "get_graphics_layout" );
            return L;
        }
        sesdf_status packGraphics( void * const * dBuffers ) const
        {
            return m_builder.check( sesdf_pack_graphics( m_builder.handle(), dBuffers ), "pack_graphics" );
        }

        sesdf::SesdfData getData()
        {
            sesdf_data d {};
            m_builder.check( sesdf_get_data( m_builder.handle(), &d ), "get_data" );
            sesdf::SesdfData out;
            out.atomNb                  = d.atom_nb;
            out.atoms                   = reinterpret_cast<float4 *>( d.atoms );
            out.convexPatchNb           = d.convex_patch_nb;
            out.convexPatches           = reinterpret_cast<uint2 *>( d.convex_patches );
            out.sectorNb                = d.sector_nb;
            out.sectors                 = reinterpret_cast<float4 *>( d.sectors );
            out.segmentPatchNb          = d.segment_patch_nb;
            out.segmentPatches          = reinterpret_cast<uint4 *>( d.segment_patches );
            out.circlePatchNb           = d.circle_patch_nb;
            out.circlePatches           = reinterpret_cast<uint2 *>( d.circle_patches );
            out.concavePatchNb          = d.concave_patch_nb;
            out.concavePatchesPosition  = reinterpret_cast<float4 *>( d.concave_patches_position );
            out.concavePatchesId        = reinterpret_cast<int4 *>( d.concave_patches_id );
            out.concavePatchesNeighbors = reinterpret_cast<float4 *>( d.concave_patches_neighbors );
            out.probeRadius             = d.probe_radius;
            out.maxConcaveNeighbors     = d.max_concave_neighbors;
            return out;
        }

        // Helpers to analyse the surface (sesdf/include/bcs/sesdf/sesdf.hpp:36-42).
        uint16_t              getMaxNeighborPerAtom() const { return uint16_t( m_builder.internal().max_neighbors ); }
        uint16_t              getMaxIntersectionsPerCircles() const { return m_builder.internal().max_intersections_per_circle; }
        uint16_t              getMaxIntersectionNeighbors() const { return m_builder.internal().max_intersection_neighbors; }
        std::vector<uint32_t> getAtomIndices() { return m_builder.atomIndices(); }
        DeviceBuffer &        getDAtomIndices() { return refresh().m_dAtomIndices; }
        DeviceBuffer &        getDAtomNeighborsCount() { return refresh().m_dNeighborsCount; }
        DeviceBuffer &        getDAtomNeighborsIndices() { return refresh().m_dNeighborsIndices; }
        DeviceBuffer &        getDCircleVisibilityStatus() { return refresh().m_dStatus; }
        // For every neighbour slot i * maxNeighbors + l (and one past the last), the number of
        // intersected circles in earlier slots: the reference's globalToTrimmed map, computed on
        // request (sesdf/include/bcs/sesdf/operations.inl:46-70).
        DeviceBuffer & getDGlobalToTrimmedId()
        {
            uint32_t * d     = nullptr;
            uint64_t   count = 0;
            m_builder.check( sesdf_get_global_to_trimmed( m_builder.handle(), &d, &count ), "get_global_to_trimmed" );
            m_dGlobalToTrimmed = DeviceBuffer( d, std::size_t( count ) * 4 );
            return m_dGlobalToTrimmed;
        }
        // Its inverse, kept by the builder: the slot of every intersected circle (trimmed id).
        DeviceBuffer & getDTrimmedToGlobalId() { return refresh().m_dTrimmedToGlobal; }

        // Counts of the last build (sizes, overflow and epsilon-band counters, rebuilds, memory).
        sesdf_counts getCounts() const { return m_builder.counts(); }
        sesdf_status lastStatus() const { return m_builder.lastStatus(); }
        sesdf_handle * handle() const { return m_builder.handle(); }

        // SDF of the built surface clamped to [-band, band] (device arrays; see sesdf_c.h).
        sesdf_status sampleGrid( const sesdf_grid_desc & grid, float band, float * dOut, bool exterior = false ) const
        {
            return m_builder.check( sesdf_sample_grid_ex( m_builder.handle(), &grid, band, exterior ? SESDF_SDF_EXTERIOR : 0, dOut ),
                                    "sample_grid" );
        }
        sesdf_status samplePoints( const float4 * dPoints, uint64_t n, float band, float * dOut, uint64_t * dCodes = nullptr,
                                   bool exterior = false ) const
        {
            return m_builder.check( sesdf_sample_points_ex( m_builder.handle(), reinterpret_cast<const float *>( dPoints ), n, band,
                                                            exterior ? SESDF_SDF_EXTERIOR : 0, dOut, dCodes ),
                                    "sample_points" );
        }
        // d(sum w_i SDF(q_i)) / d(xyz, r) per input atom; dCodes from samplePoints, or null to recompute.
        sesdf_status sdfBackward( const float4 * dPoints, const float * dWeights, uint64_t n, float band, const uint64_t * dCodes,
                                  float4 * dGradSortedScratch, float4 * dGrad, bool exterior = false ) const
        {
            return m_builder.check( sesdf_sdf_backward_ex( m_builder.handle(), reinterpret_cast<const float *>( dPoints ), dWeights, n, band,
                                                           exterior ? SESDF_SDF_EXTERIOR : 0, dCodes,
                                                           reinterpret_cast<float *>( dGradSortedScratch ), reinterpret_cast<float *>( dGrad ) ),
                                    "sdf_backward" );
        }

      private:
        Sesdf & refresh()
        {
            const sesdf_internal in = m_builder.internal();
            const sesdf_counts   c  = m_builder.counts();
            const std::size_t    n  = c.atom_nb;
            m_dAtomIndices          = DeviceBuffer( in.sorted_to_initial, n * 4 );
            m_dNeighborsCount       = DeviceBuffer( in.neighbor_count, n * 4 );
            m_dNeighborsIndices     = DeviceBuffer( in.neighbor_ids, n * in.max_neighbors * 4 );
            m_dStatus               = DeviceBuffer( in.circle_status, n * in.max_neighbors );
            m_dTrimmedToGlobal      = DeviceBuffer( in.trimmed_to_global, size_t( c.intersected_circle_nb ) * 4 );
            return *this;
        }

        detail::Builder m_builder;
        DeviceBuffer    m_dAtomIndices, m_dNeighborsCount, m_dNeighborsIndices, m_dStatus, m_dTrimmedToGlobal, m_dGlobalToTrimmed;
    };
} // namespace bcs

#endif // BCS_SESDF_SESDF_HPP
