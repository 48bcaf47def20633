/*
 * miniiso.hpp -- header-only C++ shim over the C ABI of miniiso.h, with the miniapp's class
 * interfaces, so its drivers compile against libminiiso.so unchanged apart from one alias.
 *
 *   FlyingEdgesAlgorithm(image, isoval); pass1() .. pass4(); moveOutput()
 *       flyingEdges/serial/FlyingEdgesAlgorithm.h:27-49 (OpenMP: flyingEdges/openmp/...h:29-51)
 *   MarchingCubes(image, isoval) / MarchingCubes(image, isoval, nSectionsX, nSectionsY, nSectionsZ)
 *       marchingCubes/serial/main.cpp:42-44, marchingCubes/openmp/main.cpp:159-162
 *   MarchingCubesDupFree(image, isoval, nSectionsX, nSectionsY, nSectionsZ)
 *       marchingCubes/openmpDupFree/main.cpp:159-162 (the same signature, edge-sorted vertices)
 *
 * The image and mesh types are template parameters, so this header needs none of the miniapp's
 * headers: Image must provide pointer() (the float data, x fastest), x/y/zdimension(),
 * getSpacing() and getZeroPos(); Mesh must be constructible from three vectors (points,
 * normals, std::array<size_t, 3> triangles). Errors of the C ABI become miniiso::Error
 * (a std::runtime_error). Each passN() waits for its kernels, so per-pass timers around the
 * calls measure the pass.
 */
#ifndef MINIISO_HPP
#define MINIISO_HPP

#include <array>
#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "miniiso.h"

namespace miniiso {

class Error : public std::runtime_error {
public:
    Error(int code, const std::string& what) : std::runtime_error(what), code_(code) {}
    int code() const { return code_; }

private:
    int code_;
};

inline void check(int rc, const char* what)
{
    if (rc != MISO_OK) throw Error(rc, std::string(what) + ": " + miso_last_error());
}

template <class Image>
miso_volume host_volume(Image const& image)
{
    miso_volume v{};
    v.data = image.pointer();
    v.nx = (int64_t)image.xdimension();
    v.ny = (int64_t)image.ydimension();
    v.nz = (int64_t)image.zdimension();
    const auto sp = image.getSpacing();
    const auto org = image.getZeroPos();
    for (int a = 0; a < 3; ++a) {
        v.spacing[a] = (float)sp[a];
        v.origin[a] = (float)org[a];
    }
    v.on_device = 0;
    return v;
}

/* host copy of a finished extraction into the miniapp's vector-of-arrays layout */
template <class Mesh, class Counts, class Copy>
Mesh collect(Counts counts, Copy copy)
{
    uint64_t nv = 0, nt = 0;
    check(counts(&nv, &nt), "counts");
    std::vector<std::array<float, 3> > points(nv), normals(nv);
    std::vector<std::array<size_t, 3> > tris(nt);
    static_assert(sizeof(std::array<size_t, 3>) == 3 * sizeof(uint64_t), "size_t triangles");
    check(copy(nv ? points.data()->data() : nullptr, nv ? normals.data()->data() : nullptr,
               nt ? reinterpret_cast<uint64_t*>(tris.data()->data()) : nullptr),
          "copy_out");
    return Mesh(std::move(points), std::move(normals), std::move(tris));
}

template <class Image, class Mesh>
class BasicFlyingEdges {
public:
    BasicFlyingEdges(Image const& image, float const& isoval, void* stream = nullptr)
    {
        const miso_volume v = host_volume(image);
        check(miso_fe_create(&v, isoval, stream, &h_), "FlyingEdgesAlgorithm");
    }
    ~BasicFlyingEdges() { miso_fe_destroy(h_); }
    BasicFlyingEdges(const BasicFlyingEdges&) = delete;
    BasicFlyingEdges& operator=(const BasicFlyingEdges&) = delete;

    void pass1() { check(miso_fe_pass(h_, 1), "pass1"); }
    void pass2() { check(miso_fe_pass(h_, 2), "pass2"); }
    void pass3() { check(miso_fe_pass(h_, 3), "pass3"); }
    void pass4() { check(miso_fe_pass(h_, 4), "pass4"); }

    Mesh moveOutput()
    {
        return collect<Mesh>([&](uint64_t* v, uint64_t* t) { return miso_fe_counts(h_, v, t); },
                             [&](float* p, float* n, uint64_t* t) { return miso_fe_copy_out(h_, p, n, t); });
    }

    miso_fe* handle() const { return h_; }

private:
    miso_fe* h_ = nullptr;
};

template <class Mesh, class Image, class T>
Mesh MarchingCubes(Image const& image, T const& isoval, int order = MISO_MC_ORDER_FIRST_VISIT)
{
    const miso_volume v = host_volume(image);
    miso_mc* h = nullptr;
    check(miso_mc_create(&v, (float)isoval, nullptr, &h), "MarchingCubes");
    struct Guard {
        miso_mc* h;
        ~Guard() { miso_mc_destroy(h); }
    } guard{h};
    if (order != MISO_MC_ORDER_FIRST_VISIT) check(miso_mc_set_order(h, order), "MarchingCubes order");
    check(miso_mc_run(h), "MarchingCubes run");
    return collect<Mesh>([&](uint64_t* a, uint64_t* b) { return miso_mc_counts(h, a, b); },
                         [&](float* p, float* n, uint64_t* t) { return miso_mc_copy_out(h, p, n, t); });
}

/* the OpenMP overload: the section counts only split the CPU work in the miniapp */
template <class Mesh, class Image, class T>
Mesh MarchingCubes(Image const& image, T const& isoval, unsigned, unsigned, unsigned)
{
    return MarchingCubes<Mesh>(image, isoval);
}

/* the duplicate-free OpenMP variant (marchingCubes/openmpDupFree/main.cpp:159-331): its mesh,
 * vertices sorted by global edge id as util::duplicateRemover leaves them */
template <class Mesh, class Image, class T>
Mesh MarchingCubesDupFree(Image const& image, T const& isoval, unsigned, unsigned, unsigned)
{
    return MarchingCubes<Mesh>(image, isoval, MISO_MC_ORDER_EDGE);
}

}  // namespace miniiso

#endif
