/*
 * flying_edges.cpp -- command-line driver for Flying Edges on the GPU, with the miniapp's flags,
 * report and output (flyingEdges/serial/main.cpp:27-205; the OpenMP driver differs only in the
 * algorithm name it reports).
 *
 *   miniiso_flying_edges -i volume.vtk -o mesh.vtk -v isoval [-y report.yaml]
 *
 * Missing required flags print the miniapp's message and exit with status 0, as it does.
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

using FlyingEdgesAlgorithm = miniiso::BasicFlyingEdges<Image, Mesh>;

int main(int argc, char* argv[])
{
    float isoval = 0.0f;
    bool isoval_set = false;
    const char* in = nullptr;
    const char* out = nullptr;
    std::string yaml;
    for (int i = 0; i < argc; i++) {
        const bool more = i + 1 < argc;
        if ((!strcmp(argv[i], "-i") || !strcmp(argv[i], "-input_file")) && more) {
            in = argv[++i];
        } else if ((!strcmp(argv[i], "-o") || !strcmp(argv[i], "-output_file")) && more) {
            out = argv[++i];
        } else if ((!strcmp(argv[i], "-v") || !strcmp(argv[i], "-isoval")) && more) {
            isoval_set = true;
            isoval = (float)std::atof(argv[++i]);
        } else if ((!strcmp(argv[i], "-y") || !strcmp(argv[i], "-yaml_output_file")) && more) {
            yaml = argv[++i];
        } else if (!strcmp(argv[i], "-h") || !strcmp(argv[i], "-help")) {
            std::cout << "Flying Edges (MI355X) Options:\n"
                         "  -input_file (-i)\n"
                         "  -output_file (-o)\n"
                         "  -isoval (-v)\n"
                         "  -yaml_output_file (-y)\n"
                         "  -help (-h)\n";
            return 0;
        }
    }
    if (!isoval_set || !in || !out) {
        std::cout << "Error: isoval, input_file and output_file must be set.\nTry -help\n";
        return 0;
    }
    try {
        YamlReport doc("Flying Edges", "0.1", yaml);
        doc.add("Flying Edges Algorithm", "gfx950");
        doc.add("Volume image data file path", in);
        doc.add("Polygonal mesh output file", out);
        doc.add("Isoval", isoval);
        Image image = load_vtk_image(in);
        doc.add("File x-dimension", image.xdimension());
        doc.add("File y-dimension", image.ydimension());
        doc.add("File z-dimension", image.zdimension());

        RunTimer run;
        FlyingEdgesAlgorithm algo(image, isoval);
        RunTimer t1;
        algo.pass1();
        t1.stop();
        RunTimer t2;
        algo.pass2();
        t2.stop();
        RunTimer t3;
        algo.pass3();
        t3.stop();
        RunTimer t4;
        algo.pass4();
        t4.stop();
        Mesh mesh = algo.moveOutput();
        run.stop();

        doc.add("Number of vertices in mesh", mesh.numberOfVertices());
        doc.add("Number of triangles in mesh", mesh.numberOfTriangles());
        const RunTimer* passes[4] = {&t1, &t2, &t3, &t4};
        for (int p = 0; p < 4; ++p) {
            const std::string key = "Pass " + std::to_string(p + 1);
            YamlNode* n = doc.add(key, "");
            n->add("CPU Time (clicks)", passes[p]->ticks());
            n->add("CPU Time (seconds)", passes[p]->cpu_seconds());
            n->add("Wall Time (seconds)", passes[p]->wall_seconds());
        }
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
