/*
 * vtk_legacy.hpp -- the miniapp's on-disk formats for the command-line tools.
 *
 * Volumes: VTK legacy STRUCTURED_POINTS, BINARY, big-endian scalars of type char, short,
 * unsigned_short, int, float or double converted to float; DIMENSIONS / SPACING / ORIGIN in any
 * order (reader: flyingEdges/util/LoadImage.h:27-198, ConvertBuffer.h:24-72, TypeInfo.h:23-116;
 * errors as util::file_not_found / util::bad_format, flyingEdges/util/Errors.h:26-48).
 * The Marching Cubes driver's -input_dat files: three 16-bit little-endian dimensions, then
 * nx*ny*nz 16-bit little-endian samples, unit spacing, zero origin
 * (marchingCubes/util/LoadImage.h:179-210).
 * Meshes: POINTS (big-endian float), POLYGONS as big-endian 8-byte (3, a, b, c), NORMALS
 * (writer: flyingEdges/util/SaveTriangleMesh.h:33-118).
 */
#ifndef MINIISO_APPS_VTK_LEGACY_HPP
#define MINIISO_APPS_VTK_LEGACY_HPP

#include <array>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace miso_apps {

struct FileNotFound : std::runtime_error {
    explicit FileNotFound(const std::string& f) : std::runtime_error("file not found: " + f) {}
};
struct BadFormat : std::runtime_error {
    explicit BadFormat(const std::string& m) : std::runtime_error("bad format: " + m) {}
};

/* A volume with the accessors miniiso.hpp's shim reads (util::Image3D's). */
struct Image {
    std::vector<float> data;
    std::array<size_t, 3> dims{0, 0, 0};
    std::array<float, 3> spacing{1, 1, 1};
    std::array<float, 3> origin{0, 0, 0};

    const float* pointer() const { return data.data(); }
    size_t xdimension() const { return dims[0]; }
    size_t ydimension() const { return dims[1]; }
    size_t zdimension() const { return dims[2]; }
    std::array<float, 3> getSpacing() const { return spacing; }
    std::array<float, 3> getZeroPos() const { return origin; }
};

/* util::TriangleMesh's layout and constructor. */
struct Mesh {
    std::vector<std::array<float, 3> > points, normals;
    std::vector<std::array<size_t, 3> > triangles;

    Mesh() = default;
    Mesh(std::vector<std::array<float, 3> >&& p, std::vector<std::array<float, 3> >&& n,
         std::vector<std::array<size_t, 3> >&& t)
        : points(std::move(p)), normals(std::move(n)), triangles(std::move(t))
    {
    }
    size_t numberOfVertices() const { return points.size(); }
    size_t numberOfTriangles() const { return triangles.size(); }
};

template <typename T>
T from_big_endian(const unsigned char* p)
{
    unsigned char b[sizeof(T)];
    for (size_t q = 0; q < sizeof(T); ++q) b[q] = p[sizeof(T) - 1 - q];
    T v;
    std::memcpy(&v, b, sizeof(T));
    return v;
}

template <typename T>
void to_big_endian(T v, unsigned char* p)
{
    unsigned char b[sizeof(T)];
    std::memcpy(b, &v, sizeof(T));
    for (size_t q = 0; q < sizeof(T); ++q) p[q] = b[sizeof(T) - 1 - q];
}

inline std::vector<std::string> tokens(const std::string& line)
{
    std::istringstream s(line);
    std::vector<std::string> out;
    std::string t;
    while (s >> t) out.push_back(t);
    return out;
}

inline Image load_vtk_image(const std::string& path)
{
    std::ifstream f(path, std::ios::binary);
    if (!f) throw FileNotFound(path);
    std::string line;
    std::getline(f, line);   // "# vtk DataFile Version x.y"
    std::getline(f, line);   // title
    std::getline(f, line);
    if (line != "BINARY") throw BadFormat("Only 'BINARY' format supported");
    std::getline(f, line);
    auto t = tokens(line);
    if (t.size() < 2 || t[0] != "DATASET" || t[1] != "STRUCTURED_POINTS")
        throw BadFormat("Expecting STRUCTURED_POINTS dataset");
    Image img;
    bool have[3] = {false, false, false};
    for (int q = 0; q < 3; ++q) {
        std::getline(f, line);
        t = tokens(line);
        if (t.size() < 4) throw BadFormat("Expecting DIMENSIONS, SPACING and ORIGIN");
        try {
            if (t[0] == "DIMENSIONS") {
                for (int a = 0; a < 3; ++a) img.dims[a] = std::stoull(t[1 + a]);
                have[0] = true;
            } else if (t[0] == "SPACING") {
                for (int a = 0; a < 3; ++a) img.spacing[a] = std::stof(t[1 + a]);
                have[1] = true;
            } else if (t[0] == "ORIGIN") {
                for (int a = 0; a < 3; ++a) img.origin[a] = std::stof(t[1 + a]);
                have[2] = true;
            } else {
                throw BadFormat("Expecting DIMENSIONS, SPACING and ORIGIN");
            }
        } catch (const std::logic_error&) {
            throw BadFormat("Expecting " + t[0] + " [3]");
        }
    }
    if (!(have[0] && have[1] && have[2])) throw BadFormat("Expecting DIMENSIONS, SPACING and ORIGIN");
    std::getline(f, line);
    t = tokens(line);
    if (t.size() < 2 || t[0] != "POINT_DATA") throw BadFormat("Expecting POINT_DATA <npoints>");
    const size_t npoints = std::stoull(t[1]);
    std::getline(f, line);
    t = tokens(line);
    if (t.size() < 3 || t[0] != "SCALARS") throw BadFormat("Expecting SCALARS <name> <type>");
    const std::string type = t[2];
    std::getline(f, line);
    t = tokens(line);
    if (t.empty() || t[0] != "LOOKUP_TABLE") throw BadFormat("Expecting LOOKUP_TABLE name");
    size_t width = 0;
    if (type == "char") width = 1;
    else if (type == "short" || type == "unsigned_short") width = 2;
    else if (type == "int" || type == "float") width = 4;
    else if (type == "double") width = 8;
    else throw BadFormat("Unsupported data type");
    std::vector<unsigned char> raw(npoints * width);
    f.read(reinterpret_cast<char*>(raw.data()), (std::streamsize)raw.size());
    const size_t got = (size_t)f.gcount() / width;
    const size_t n = img.dims[0] * img.dims[1] * img.dims[2];
    img.data.assign(n, 0.0f);
    for (size_t q = 0; q < std::min(got, n); ++q) {
        const unsigned char* p = raw.data() + q * width;
        float v = 0.0f;
        if (type == "char") v = (float)(signed char)p[0];
        else if (type == "short") v = (float)from_big_endian<int16_t>(p);
        else if (type == "unsigned_short") v = (float)from_big_endian<uint16_t>(p);
        else if (type == "int") v = (float)from_big_endian<int32_t>(p);
        else if (type == "float") v = from_big_endian<float>(p);
        else v = (float)from_big_endian<double>(p);
        img.data[q] = v;
    }
    return img;
}

inline Image load_dat_image(const std::string& path)
{
    std::ifstream f(path, std::ios::binary);
    if (!f) throw FileNotFound(path);
    Image img;
    unsigned char b[2];
    for (int a = 0; a < 3; ++a) {
        f.read(reinterpret_cast<char*>(b), 2);
        img.dims[a] = (size_t)b[0] | (size_t)b[1] << 8;
    }
    const size_t n = img.dims[0] * img.dims[1] * img.dims[2];
    img.data.assign(n, 0.0f);
    for (size_t q = 0; q < n; ++q) {
        if (!f.read(reinterpret_cast<char*>(b), 2)) break;
        img.data[q] = (float)((unsigned)b[0] | (unsigned)b[1] << 8);
    }
    return img;
}

inline void save_vtk_mesh(const Mesh& m, const std::string& path)
{
    std::ofstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("cannot write " + path);
    const size_t v = m.points.size(), t = m.triangles.size();
    f << "# vtk DataFile Version 3.0\nIsosurface Mesh\nBINARY\nDATASET POLYDATA\n";
    f << "POINTS " << v << " float\n";
    std::vector<unsigned char> buf(12 * v);
    for (size_t q = 0; q < v; ++q)
        for (int c = 0; c < 3; ++c) to_big_endian<float>(m.points[q][c], &buf[12 * q + 4 * c]);
    f.write(reinterpret_cast<const char*>(buf.data()), (std::streamsize)buf.size());
    f << "\nPOLYGONS " << t << " " << 4 * t << "\n";
    buf.assign(32 * t, 0);
    for (size_t q = 0; q < t; ++q) {
        to_big_endian<uint64_t>(3, &buf[32 * q]);
        for (int c = 0; c < 3; ++c) to_big_endian<uint64_t>((uint64_t)m.triangles[q][c], &buf[32 * q + 8 * (c + 1)]);
    }
    f.write(reinterpret_cast<const char*>(buf.data()), (std::streamsize)buf.size());
    f << "\nPOINT_DATA " << v << "\nNORMALS Normals float\n";
    buf.assign(12 * v, 0);
    for (size_t q = 0; q < v; ++q)
        for (int c = 0; c < 3; ++c) to_big_endian<float>(m.normals[q][c], &buf[12 * q + 4 * c]);
    f.write(reinterpret_cast<const char*>(buf.data()), (std::streamsize)buf.size());
    f << "\n";
}

inline Mesh load_vtk_mesh(const std::string& path)
{
    std::ifstream f(path, std::ios::binary);
    if (!f) throw FileNotFound(path);
    std::string line;
    for (int q = 0; q < 4; ++q) std::getline(f, line);
    std::getline(f, line);
    auto t = tokens(line);
    if (t.size() < 2 || t[0] != "POINTS") throw BadFormat("Expecting POINTS <n> float");
    const size_t v = std::stoull(t[1]);
    Mesh m;
    m.points.resize(v);
    m.normals.resize(v);
    std::vector<unsigned char> buf(12 * v);
    f.read(reinterpret_cast<char*>(buf.data()), (std::streamsize)buf.size());
    for (size_t q = 0; q < v; ++q)
        for (int c = 0; c < 3; ++c) m.points[q][c] = from_big_endian<float>(&buf[12 * q + 4 * c]);
    std::getline(f, line);
    std::getline(f, line);
    t = tokens(line);
    if (t.size() < 2 || t[0] != "POLYGONS") throw BadFormat("Expecting POLYGONS <n> <4n>");
    const size_t nt = std::stoull(t[1]);
    m.triangles.resize(nt);
    buf.resize(32 * nt);
    f.read(reinterpret_cast<char*>(buf.data()), (std::streamsize)buf.size());
    for (size_t q = 0; q < nt; ++q)
        for (int c = 0; c < 3; ++c) m.triangles[q][c] = (size_t)from_big_endian<uint64_t>(&buf[32 * q + 8 * (c + 1)]);
    std::getline(f, line);
    std::getline(f, line);
    std::getline(f, line);
    buf.resize(12 * v);
    f.read(reinterpret_cast<char*>(buf.data()), (std::streamsize)buf.size());
    for (size_t q = 0; q < v; ++q)
        for (int c = 0; c < 3; ++c) m.normals[q][c] = from_big_endian<float>(&buf[12 * q + 4 * c]);
    return m;
}

}  // namespace miso_apps

#endif
