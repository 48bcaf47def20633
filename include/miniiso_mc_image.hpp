/*
 * miniiso_mc_image.hpp -- binds the Marching Cubes miniapp's own driver to libminiiso.so.
 *
 * The miniapp defines MarchingCubes<T>(util::Image3D<T> const&, T const&) in its driver's
 * translation unit (marchingCubes/serial/main.cpp:42-192) and its image class
 * (marchingCubes/util/Image3D.h:26-156) keeps the voxel vector, spacing and origin private, with
 * no accessor. This header, included after the miniapp's Image3D.h and TriangleMesh.h, supplies
 *
 *   MCImageView            the Image interface of miniiso.hpp (pointer(), x/y/zdimension(),
 *                          getSpacing(), getZeroPos()) over a whole-volume util::Image3D<float>;
 *   ::MarchingCubes(image, isoval)
 *                          the miniapp's signature, running the extraction on the GPU with the
 *                          serial reference's first-visit vertex order;
 *
 * so that the driver compiles unchanged once its own definition is renamed out of the way
 * (oracle/Makefile, target `integration`, does that on the fly). The private members are read
 * through explicit template instantiations, where C++ does not apply access checking to the
 * names used ([temp.explicit]); nothing of the miniapp is edited or copied. Sectioned images
 * (the MPI variant's, marchingCubes/util/Image3D.h:36-49) are refused: only whole volumes carry
 * every voxel.
 */
#ifndef MINIISO_MC_IMAGE_HPP
#define MINIISO_MC_IMAGE_HPP

#include <array>
#include <cstddef>
#include <vector>

#include "miniiso.hpp"

namespace miniiso {
namespace mc_detail {

template <class Tag, typename Tag::type M>
struct Expose {
    friend typename Tag::type member(Tag) { return M; }
};

using McImage = util::Image3D<float>;

struct DataTag {
    using type = std::vector<float> McImage::*;
    friend type member(DataTag);
};
struct SpacingTag {
    using type = std::array<float, 3> McImage::*;
    friend type member(SpacingTag);
};
struct ZeroPosTag {
    using type = std::array<float, 3> McImage::*;
    friend type member(ZeroPosTag);
};
struct DataBegTag {
    using type = std::array<std::size_t, 3> McImage::*;
    friend type member(DataBegTag);
};
struct DataEndTag {
    using type = std::array<std::size_t, 3> McImage::*;
    friend type member(DataEndTag);
};

template struct Expose<DataTag, &McImage::data>;
template struct Expose<SpacingTag, &McImage::spacing>;
template struct Expose<ZeroPosTag, &McImage::zeroPos>;
template struct Expose<DataBegTag, &McImage::dataBeg>;
template struct Expose<DataEndTag, &McImage::dataEnd>;

}  // namespace mc_detail

/* The Image interface of miniiso.hpp over the Marching Cubes miniapp's whole-volume image. */
class MCImageView {
public:
    explicit MCImageView(mc_detail::McImage const& image) : im_(image)
    {
        const auto& beg = im_.*member(mc_detail::DataBegTag{});
        const auto& end = im_.*member(mc_detail::DataEndTag{});
        const std::array<std::size_t, 3> dims{im_.xdimension(), im_.ydimension(), im_.zdimension()};
        if (beg[0] || beg[1] || beg[2] || end != dims || (im_.*member(mc_detail::DataTag{})).size() !=
                                                             dims[0] * dims[1] * dims[2])
            throw Error(MISO_ERR_ARG, "MarchingCubes: the image must hold the whole volume (no sections)");
    }
    const float* pointer() const { return (im_.*member(mc_detail::DataTag{})).data(); }
    std::size_t xdimension() const { return im_.xdimension(); }
    std::size_t ydimension() const { return im_.ydimension(); }
    std::size_t zdimension() const { return im_.zdimension(); }
    std::array<float, 3> getSpacing() const { return im_.*member(mc_detail::SpacingTag{}); }
    std::array<float, 3> getZeroPos() const { return im_.*member(mc_detail::ZeroPosTag{}); }

private:
    mc_detail::McImage const& im_;
};

}  // namespace miniiso

/* MarchingCubes(image, isoval), the signature of marchingCubes/serial/main.cpp:42-44, on the GPU:
 * the same mesh as the serial reference, vertices in its first-visit order. */
template <typename T>
util::TriangleMesh<T> MarchingCubes(util::Image3D<T> const& image, T const& isoval)
{
    return miniiso::MarchingCubes<util::TriangleMesh<T> >(miniiso::MCImageView(image), isoval);
}

#endif
