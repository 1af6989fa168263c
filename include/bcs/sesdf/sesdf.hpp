// bcs::Sesdf: complete solvent-excluded surface of one molecule on an MI355X, over the C ABI of
// libsesdf_hip.so. Same constructor, build() and getters as the reference class
// (sesdf/include/bcs/sesdf/sesdf.hpp:15-88); header-only, link with -lsesdf_hip.
//
// Differences from the reference, by design:
//   - graphics = true does not create OpenGL buffers (no GL interop): getGraphics() returns sizes
//     and offsets with zero ids, getData() returns the device pointers;
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

        sesdf::SesdfGraphics getGraphics() const
        {
            sesdf_data d {};
            m_builder.check( sesdf_get_data( m_builder.handle(), &d ), "get_data" );
            sesdf::SesdfGraphics g;
            auto span = []( std::size_t bytes ) {
                HandleSpan<GLuint> h;
                h.size = bytes;
                return h;
            };
            g.atomNb                  = d.atom_nb;
            g.atoms                   = span( size_t( d.atom_nb ) * 16 );
            g.convexPatchNb           = d.convex_patch_nb;
            g.convexPatches           = span( size_t( d.convex_patch_nb ) * 8 );
            g.sectorNb                = d.sector_nb;
            g.sectors                 = span( size_t( d.sector_nb ) * 16 );
            g.segmentPatchNb          = d.segment_patch_nb;
            g.segmentPatches          = span( size_t( d.segment_patch_nb ) * 16 );
            g.circlePatchNb           = d.circle_patch_nb;
            g.circlePatches           = span( size_t( d.circle_patch_nb ) * 8 );
            g.concavePatchNb          = d.concave_patch_nb;
            g.concavePatchesPosition  = span( size_t( d.concave_patch_nb ) * 16 );
            g.concavePatchesId        = span( size_t( d.concave_patch_nb ) * 16 );
            const sesdf_counts c      = m_builder.counts();
            g.concavePatchesNeighbors = span( size_t( c.concave_with_neighbors_nb ) * d.max_concave_neighbors * 16 );
            g.probeRadius             = d.probe_radius;
            g.maxConcaveNeighbors     = d.max_concave_neighbors;
            return g;
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
        // The builder keeps the inverse of the reference's globalToTrimmed map: for every intersected
        // circle (trimmed id) its global slot i * maxNeighbors + l.
        DeviceBuffer &        getDGlobalToTrimmedId() { return refresh().m_dTrimmedToGlobal; }

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
        DeviceBuffer    m_dAtomIndices, m_dNeighborsCount, m_dNeighborsIndices, m_dStatus, m_dTrimmedToGlobal;
    };
} // namespace bcs

#endif // BCS_SESDF_SESDF_HPP
