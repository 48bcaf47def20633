/*
 * same_contents.cpp -- order-independent mesh equivalence of two mesh files, with the miniapp
 * checker's rule and messages (utilities/tests/SameContentsCheck.cpp:190-349):
 *   - the triangle counts are equal;
 *   - every point of each mesh occurs among the other's points, and likewise every normal
 *     (floats compared with ==, so -0 equals 0 and a NaN matches nothing);
 *   - every triangle of the first, as its three corner positions in order, occurs among the
 *     second's triangles.
 * Vertex counts and vertex numbering may differ (a mesh may repeat points).
 *
 *   miniiso_same_contents a.vtk b.vtk      prints "The two meshes are equivalent." or
 *                                          "ERROR: The two meshes are not equivalent."
 * The exit status is 0 in both cases, as the miniapp's checker; 2 if a file cannot be read.
 */
#include <algorithm>
#include <array>
#include <cstring>
#include <iostream>
#include <vector>

#include "vtk_legacy.hpp"

using namespace miso_apps;

namespace {

/* bit pattern of a float under == equality: -0 -> +0; NaN flagged */
inline uint32_t key_bits(float f, bool* nan)
{
    if (f != f) *nan = true;
    if (f == 0.0f) f = 0.0f;
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return u;
}

template <size_t N>
std::vector<std::array<uint32_t, N> > sorted_keys(const std::vector<std::array<float, N> >& v, bool* nan)
{
    std::vector<std::array<uint32_t, N> > k(v.size());
    for (size_t q = 0; q < v.size(); ++q)
        for (size_t c = 0; c < N; ++c) k[q][c] = key_bits(v[q][c], nan);
    std::sort(k.begin(), k.end());
    return k;
}

/* every element of a occurs in b (both sorted) */
template <class K>
bool contained(const std::vector<K>& a, const std::vector<K>& b)
{
    for (const K& x : a)
        if (!std::binary_search(b.begin(), b.end(), x)) return false;
    return true;
}

std::vector<std::array<float, 9> > corner_triangles(const Mesh& m)
{
    std::vector<std::array<float, 9> > t(m.triangles.size());
    for (size_t q = 0; q < t.size(); ++q)
        for (int c = 0; c < 3; ++c)
            for (int d = 0; d < 3; ++d) t[q][3 * c + d] = m.points.at(m.triangles[q][c])[d];
    return t;
}

const char* compare(const Mesh& a, const Mesh& b)
{
    if (a.numberOfTriangles() != b.numberOfTriangles()) return "not the same number of triangles";
    bool pnan = false, nnan = false, tnan = false;
    const auto pa = sorted_keys(a.points, &pnan), pb = sorted_keys(b.points, &pnan);
    const auto na = sorted_keys(a.normals, &nnan), nb = sorted_keys(b.normals, &nnan);
    if (pnan || !contained(pa, pb) || !contained(pb, pa)) return "point not found";
    if (nnan || !contained(na, nb) || !contained(nb, na)) return "normal not found";
    const auto ta = sorted_keys(corner_triangles(a), &tnan), tb = sorted_keys(corner_triangles(b), &tnan);
    if (!contained(ta, tb)) return "triangle not found";
    return nullptr;
}

}  // namespace

int main(int argc, char* argv[])
{
    if (argc < 3) {
        std::cout << "usage: miniiso_same_contents meshA.vtk meshB.vtk\n";
        return 2;
    }
    try {
        const Mesh a = load_vtk_mesh(argv[1]);
        const Mesh b = load_vtk_mesh(argv[2]);
        const char* why = compare(a, b);
        if (why) {
            if (!std::strcmp(why, "not the same number of triangles"))
                std::cout << a.numberOfTriangles() << ", " << b.numberOfTriangles() << "\n";
            std::cout << why << "\n";
            std::cout << "ERROR: The two meshes are not equivalent." << std::endl;
        } else {
            std::cout << "The two meshes are equivalent." << std::endl;
        }
    } catch (const std::exception& e) {
        std::cerr << e.what() << "\n";
        return 2;
    }
    return 0;
}
