/*
 * tables.h -- marching-cubes case data for the device kernels and the host code.
 *
 * The triangulation of each of the 256 cube cases must be the reference's
 * (flyingEdges/util/MarchingCubesTables.h:303-561; the Marching Cubes copy in
 * marchingCubes/util/MarchingCubesTables.h is identical) because triangle order and
 * connectivity are compared bit-exactly. Here a case is one 64-bit word: cube-edge ids as
 * 4-bit nibbles, three per triangle, lowest nibble first, padded with 0xF. One load yields a
 * whole case, and the word fits a register or 2 KB of LDS for the full table. numTris and isCut
 * (reference :23-301) are derived from it: a case cuts exactly the edges its triangles name.
 *
 * Cube numbering (flyingEdges/serial/FlyingEdgesAlgorithm.cpp:138-157): corner q of cell (i,j,k)
 * is bit q of the case, v0=(i,j,k) v1=(i+1,j,k) v2=(i+1,j+1,k) v3=(i,j+1,k), v4..v7 the same at
 * k+1. Edge e runs from corner kEdgeLo[e] to kEdgeHi[e], always toward the higher coordinate
 * (reference edgeVertices, MarchingCubesTables.h:563-566).
 */
#ifndef MINIISO_TABLES_H
#define MINIISO_TABLES_H

#include <stdint.h>

#if defined(__HIPCC__)
#define MISO_HD __host__ __device__
#else
#define MISO_HD
#endif

namespace miso {

static constexpr uint64_t kCaseEdges[256] = {
    0xffffffffffffffffull, 0xfffffffffffff830ull, 0xfffffffffffff190ull, 0xffffffffff819831ull,
    0xfffffffffffff2b1ull, 0xffffffffff2b1830ull, 0xffffffffff2902b9ull, 0xfffffff98bb82832ull,
    0xfffffffffffffa23ull, 0xffffffffffa08a20ull, 0xffffffffff3a2901ull, 0xfffffff8a99a1a21ull,
    0xffffffffffb3ab13ull, 0xfffffffab88b0b10ull, 0xfffffffb9aa93903ull, 0xffffffffff8ab8b9ull,
    0xfffffffffffff784ull, 0xffffffffff347304ull, 0xffffffffff478190ull, 0xfffffff317714194ull,
    0xffffffffff4782b1ull, 0xfffffff2b1043473ull, 0xfffffff4780292b9ull, 0xffff947732972b92ull,
    0xffffffffffa23478ull, 0xfffffff04224a47aull, 0xfffffff3a2478019ull, 0xffff219a294a97a4ull,
    0xfffffff847ab3b13ull, 0xffffa470414a1ab1ull, 0xffff03aab90a9784ull, 0xfffffffab9a947a4ull,
    0xfffffffffffff549ull, 0xffffffffff830549ull, 0xffffffffff501540ull, 0xfffffff153358548ull,
    0xffffffffff5492b1ull, 0xfffffff9542b1083ull, 0xfffffff0244252b5ull, 0xffff483543253b52ull,
    0xffffffffff3a2549ull, 0xfffffff9548a0a20ull, 0xfffffff3a2150540ull, 0xffff8548a2582152ull,
    0xfffffff54913b3abull, 0xffffab8b18810954ull, 0xffff03aab50a5405ull, 0xfffffff8ab8b5485ull,
    0xffffffffff795789ull, 0xfffffff735539309ull, 0xfffffff571170780ull, 0xffffffffff573531ull,
    0xfffffff12b579789ull, 0xffff73530550912bull, 0xffff52b578258028ull, 0xfffffff573532b52ull,
    0xfffffffa23897957ull, 0xffff7a2209729579ull, 0xffff5717811803a2ull, 0xfffffff15717a21aull,
    0xffff3ab13b578589ull, 0xfb0a0b1a07095705ull, 0xf70507850b03ab0aull, 0xffffffffffa57b5aull,
    0xfffffffffffff65bull, 0xffffffffffb65830ull, 0xffffffffffb65019ull, 0xfffffffb65981831ull,
    0xffffffffff612651ull, 0xfffffff083261651ull, 0xfffffff260069659ull, 0xffff283265825985ull,
    0xffffffffff65b3a2ull, 0xfffffff65b20a08aull, 0xfffffffb653a2190ull, 0xffff8a9a29921b65ull,
    0xfffffff1355363a6ull, 0xffffa65510a508a0ull, 0xffff590650360a63ull, 0xfffffff98a9a6596ull,
    0xffffffffff784b65ull, 0xfffffff5b6734304ull, 0xfffffff478b65901ull, 0xffff94773197165bull,
    0xfffffff784516126ull, 0xffff473043265251ull, 0xffff260650059478ull, 0xf692965293947397ull,
    0xfffffff65b847a23ull, 0xffff7a2204724b65ull, 0xffffb653a2784190ull, 0xfb65a474a9a29219ull,
    0xffffa65513a53478ull, 0xf4a0a470a1a651a5ull, 0xf47863a360650590ull, 0xffffa977949a6596ull,
    0xffffffffff4b649bull, 0xfffffff8309b4b64ull, 0xfffffff40660b01bull, 0xffff1b6648168318ull,
    0xfffffff642241491ull, 0xffff642492291083ull, 0xffffffffff264240ull, 0xfffffff264248328ull,
    0xfffffff23a64b49bull, 0xffffb649b48a2820ull, 0xffff1b6640160a23ull, 0xfa181a28141b6416ull,
    0xffff63a139369649ull, 0xf41614961a108a18ull, 0xfffffff640603a63ull, 0xffffffffff68a486ull,
    0xfffffff9b88b7b67ull, 0xffff7b69b0b70730ull, 0xffff801781b7167bull, 0xfffffff73171b67bull,
    0xffff678891681261ull, 0xf397930796912692ull, 0xfffffff026067807ull, 0xffffffffff726327ull,
    0xffff67889b68b3a2ull, 0xfb797b69707a2072ull, 0xf3a27b6b71781801ull, 0xffff71661b17a21aull,
    0xf36163a169678968ull, 0xffffffffff67a910ull, 0xffff60aa03067807ull, 0xfffffffffffffa67ull,
    0xfffffffffffff6a7ull, 0xffffffffff76a083ull, 0xffffffffff76a190ull, 0xfffffff76a318198ull,
    0xffffffffffa7612bull, 0xfffffffa760832b1ull, 0xfffffffa76b92902ull, 0xffff98b83bb32a76ull,
    0xffffffffff276237ull, 0xfffffff206607087ull, 0xfffffff190372762ull, 0xffff768981861621ull,
    0xfffffff37117b76bull, 0xffff0818717b176bull, 0xffffb76b907b0370ull, 0xfffffffb98b876b7ull,
    0xffffffffff86a846ull, 0xfffffff4600636a3ull, 0xfffffff0194686a8ull, 0xffff36a319639469ull,
    0xfffffffb12a86846ull, 0xffff4606a00a32b1ull, 0xffffb922906a4a84ull, 0xf63436a43932b93bull,
    0xfffffff624428238ull, 0xffffffffff624420ull, 0xffff384462342901ull, 0xfffffff462421941ull,
    0xffffb16468618138ull, 0xfffffff04606b10bull, 0xf93b390b36384634ull, 0xffffffffffb4694bull,
    0xffffffffff6a7954ull, 0xfffffff76a954830ull, 0xfffffff6a7405015ull, 0xffff15354334876aull,
    0xfffffff6a712b549ull, 0xffff9548302b1a76ull, 0xffff0242b44b56a7ull, 0xf76a52b253543483ull,
    0xfffffff495627237ull, 0xffff876620860549ull, 0xffff405501763623ull, 0xf581854182876286ull,
    0xffff37176116b549ull, 0xf5497080717616b1ull, 0xf7b3b763b0b540b4ull, 0xffff8b44b5b876b7ull,
    0xfffffff89aa96956ull, 0xffff9505606306a3ull, 0xffff6a51505a0a80ull, 0xfffffff315356a36ull,
    0xffff56aa895a92b1ull, 0xf2b16959606a0a30ull, 0xf25052b05856a85aull, 0xffff53bb32356a36ull,
    0xffff823625285895ull, 0xfffffff620609569ull, 0xf286823685801581ull, 0xffffffffff162561ull,
    0xf9686958636b1361ull, 0xffff60550906b10bull, 0xffffffffff6b5380ull, 0xfffffffffffff56bull,
    0xffffffffff5a75baull, 0xfffffff30875a5baull, 0xfffffff901ba5a75ull, 0xffff318819a7b75bull,
    0xfffffff51771a12aull, 0xffff2a7751271830ull, 0xffffa72029279759ull, 0xf8292839252a7527ull,
    0xfffffff7533525b2ull, 0xffff25b758528208ull, 0xffffb23375b35019ull, 0xf52725b728219829ull,
    0xffffffffff753351ull, 0xfffffff751710870ull, 0xfffffff375359039ull, 0xffffffffff975879ull,
    0xfffffffa8bb85845ull, 0xffff30aba5a05045ull, 0xffff45bba84b8190ull, 0xf14341934a45ba4bull,
    0xffff584a82852512ull, 0xf1a5a125a4a304a0ull, 0xf85a584a52590250ull, 0xffffffffffa32459ull,
    0xffff8434535235b2ull, 0xfffffff204245b25ull, 0xf1905848535b3b23ull, 0xffff429921245b25ull,
    0xfffffff513538458ull, 0xffffffffff051450ull, 0xffff350059538458ull, 0xfffffffffffff459ull,
    0xfffffffba99a4a74ull, 0xffffba9a79974830ull, 0xffff4a7401a41ba1ull, 0xfa4b4a7b41483143ull,
    0xffff1292a9a49a74ull, 0xf830a121a9a79749ull, 0xfffffff40242a74aull, 0xffff24334842a74aull,
    0xffff4973727929b2ull, 0xf07270827b749b79ull, 0xf0b4b014b7b237b3ull, 0xffffffffff748b21ull,
    0xfffffff137174914ull, 0xffff718810174914ull, 0xffffffffff437034ull, 0xfffffffffffff874ull,
    0xffffffffffa8bb89ull, 0xfffffff9ba9a3093ull, 0xfffffffba8b801b0ull, 0xffffffffff3ba1b3ull,
    0xfffffffa89a912a1ull, 0xffffa922919a3093ull, 0xffffffffff0a82a0ull, 0xfffffffffffff2a3ull,
    0xfffffff89b8b2382ull, 0xffffffffff920b29ull, 0xffffb811808b2382ull, 0xfffffffffffffb21ull,
    0xffffffffff189381ull, 0xfffffffffffff910ull, 0xfffffffffffff380ull, 0xffffffffffffffffull,
};

static constexpr uint8_t kEdgeLo[12] = {0, 1, 3, 0, 4, 5, 7, 4, 0, 1, 3, 2};
static constexpr uint8_t kEdgeHi[12] = {1, 2, 2, 3, 5, 6, 6, 7, 4, 5, 7, 6};

/* Edge geometry relative to the cell (i,j,k): axis (0=x,1=y,2=z) and the offset of its lower
 * corner. Derived from kEdgeLo/kEdgeHi and the corner numbering above. */
static constexpr uint8_t kEdgeAxis[12] = {0, 1, 0, 1, 0, 1, 0, 1, 2, 2, 2, 2};
static constexpr uint8_t kEdgeDx[12]   = {0, 1, 0, 0, 0, 1, 0, 0, 0, 1, 0, 1};
static constexpr uint8_t kEdgeDy[12]   = {0, 0, 1, 0, 0, 0, 1, 0, 0, 0, 1, 1};
static constexpr uint8_t kEdgeDz[12]   = {0, 0, 0, 0, 1, 1, 1, 1, 0, 0, 0, 0};

MISO_HD constexpr int case_tri_count(uint64_t packed)
{
    int n = 0;
    while (n < 15 && ((packed >> (4 * n)) & 0xF) != 0xF) ++n;
    return n / 3;
}

MISO_HD constexpr int case_edge(uint64_t packed, int slot) { return (int)((packed >> (4 * slot)) & 0xF); }

}  // namespace miso

#endif
