// bcs::load from C++, as the reference's applications call it (app/benchmarks/main.cu:177,
// app/circles/main.cu:117): prints the atom count, the bounding box of the vdW spheres and the sum
// of radii of each file given, one line per file; a file that cannot be read prints its error and
// makes the exit status 1.
#include <cstdio>
#include <exception>

#include "bcs/core/loader.hpp"

int main( int argc, char ** argv )
{
    int status = 0;
    for ( int a = 1; a < argc; a++ )
    {
        try
        {
            const bcs::Molecule molecule = bcs::load( argv[ a ] );
            const bcs::Aabb     box      = bcs::getAabb( bcs::ConstSpan<bcs::Vec4f> { molecule.data(), molecule.size() } );
            double              radii    = 0.0;
            for ( const bcs::Vec4f & atom : molecule )
                radii += atom.w;
            std::printf( "%s %zu %.4f %.4f %.4f %.4f %.4f %.4f %.6f\n", argv[ a ], molecule.size(), box.min.x, box.min.y, box.min.z,
                         box.max.x, box.max.y, box.max.z, radii );
        }
        catch ( const std::exception & e )
        {
            std::printf( "%s error %s\n", argv[ a ], e.what() );
            status = 1;
        }
    }
    return status;
}
