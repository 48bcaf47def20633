/*
 * marching_cubes.cpp -- command-line driver for Marching Cubes on the GPU, with the miniapp's
 * flags, report and output (marchingCubes/serial/main.cpp:194-330), including -input_dat.
 * -dup_free writes the duplicate-free OpenMP variant's vertex order instead (vertices sorted by
 * global edge id, marchingCubes/openmpDupFree/main.cpp:324-331); that driver's section flags
 * (-sx, -sy, -sz, marchingCubes/openmpDupFree/main.cpp:376-387) are accepted and do not change
 * the result.
 *
 *   miniiso_marching_cubes -i volume.vtk | -input_dat volume.dat -o mesh.vtk -v isoval [-y report.yaml]
 *                          [-dup_free] [-sx N -sy N -sz N]
 */
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <string>

#include "miniiso.hpp"
#include "run_timer.hpp"
#include "vtk_legacy.hpp"
#include "yaml_report.hpp"

using namespace miso_apps;

int main(int argc, char* argv[])
{
    float isoval = 0.0f;
    bool isoval_set = false, use_dat = false, dup_free = false;
    const char* in = nullptr;
    const char* out = nullptr;
    std::string yaml;
    for (int i = 0; i < argc; i++) {
        const bool more = i + 1 < argc;
        if (!strcmp(argv[i], "-input_dat") && more) {
            in = argv[++i];
            use_dat = true;
        } else if ((!strcmp(argv[i], "-i") || !strcmp(argv[i], "-input_file")) && more) {
            in = argv[++i];
        } else if ((!strcmp(argv[i], "-o") || !strcmp(argv[i], "-output_file")) && more) {
            out = argv[++i];
        } else if ((!strcmp(argv[i], "-v") || !strcmp(argv[i], "-isoval")) && more) {
            isoval_set = true;
            isoval = (float)std::atof(argv[++i]);
        } else if ((!strcmp(argv[i], "-y") || !strcmp(argv[i], "-yaml_output_file")) && more) {
            yaml = argv[++i];
        } else if (!strcmp(argv[i], "-dup_free")) {
            dup_free = true;
        } else if ((!strcmp(argv[i], "-sx") || !strcmp(argv[i], "-sections_x") || !strcmp(argv[i], "-sy") ||
                    !strcmp(argv[i], "-sections_y") || !strcmp(argv[i], "-sz") || !strcmp(argv[i], "-sections_z")) &&
                   more) {
            ++i;
        } else if (!strcmp(argv[i], "-h") || !strcmp(argv[i], "-help")) {
            std::cout << "Marching Cubes (MI355X) Options:\n"
                         "  -input_file (-i)\n"
                         "  -input_dat\n"
                         "  -output_file (-o)\n"
                         "  -isoval (-v)\n"
                         "  -yaml_output_file (-y)\n"
                         "  -dup_free (vertices sorted by global edge id)\n"
                         "  -sections_x (-sx), -sections_y (-sy), -sections_z (-sz) (accepted, no effect)\n"
                         "  -help (-h)\n";
            return 0;
        }
    }
    if (!isoval_set || !in || !out) {
        std::cout << "Error: isoval, input_file and output_file must be set.\nTry -help\n";
        return 0;
    }
    try {
        YamlReport doc("Marching Cubes", "0.2", yaml);
        doc.add("Marching Cubes Algorithm", "gfx950");
        doc.add("Volume image data file path", in);
        doc.add("Polygonal mesh output file", out);
        doc.add("Isoval", isoval);
        Image image = use_dat ? load_dat_image(in) : load_vtk_image(in);
        doc.add("File x-dimension", image.xdimension());
        doc.add("File y-dimension", image.ydimension());
        doc.add("File z-dimension", image.zdimension());
        RunTimer run;
        Mesh mesh = miniiso::MarchingCubes<Mesh>(image, isoval,
                                                 dup_free ? MISO_MC_ORDER_EDGE : MISO_MC_ORDER_FIRST_VISIT);
        run.stop();
        doc.add("Number of vertices in mesh", mesh.numberOfVertices());
        doc.add("Number of triangles in mesh", mesh.numberOfTriangles());
        doc.add("Total Program CPU Time (clicks)", run.ticks());
        doc.add("Total Program CPU Time (seconds)", run.cpu_seconds());
        doc.add("Total Program WALL Time (seconds)", run.wall_seconds());
        std::cout << doc.generate();
        save_vtk_mesh(mesh, out);
    } catch (const std::exception& e) {
        std::cerr << e.what() << "\n";
        return 1;
    }
    return 0;
}
