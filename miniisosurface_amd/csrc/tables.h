/*
 * tables.h -- marching-cubes case data for the device kernels and the host code.
 *
 * The triangulation of each of the 256 cube cases must be the reference's
 * (flyingEdges/util/MarchingCubesTables.h:303-561; the Marching Cubes copy in
 * marchingCubes/util/MarchingCubesTables.h is identical) because triangle order and
 * connectivity are compared bit-exactly. Here a case is one 64-bit word: cube-edge ids as
 * 4-bit nibbles, three per triangle, lowest nibble first, padded with 0xF. One load yields a
 * whole case, and the word fits a register or 2 KB of LDS for the full table. The top nibble
 * (bits 60-63, never needed for edges: a case has at most 15) holds the triangle count, the
 * reference's numTris (:23-41); isCut (:43-301) is derived: a case cuts exactly the edges its
 * triangles name.
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
    0x0fffffffffffffffull, 0x1ffffffffffff830ull, 0x1ffffffffffff190ull, 0x2fffffffff819831ull,
    0x1ffffffffffff2b1ull, 0x2fffffffff2b1830ull, 0x2fffffffff2902b9ull, 0x3ffffff98bb82832ull,
    0x1ffffffffffffa23ull, 0x2fffffffffa08a20ull, 0x2fffffffff3a2901ull, 0x3ffffff8a99a1a21ull,
    0x2fffffffffb3ab13ull, 0x3ffffffab88b0b10ull, 0x3ffffffb9aa93903ull, 0x2fffffffff8ab8b9ull,
    0x1ffffffffffff784ull, 0x2fffffffff347304ull, 0x2fffffffff478190ull, 0x3ffffff317714194ull,
    0x2fffffffff4782b1ull, 0x3ffffff2b1043473ull, 0x3ffffff4780292b9ull, 0x4fff947732972b92ull,
    0x2fffffffffa23478ull, 0x3ffffff04224a47aull, 0x3ffffff3a2478019ull, 0x4fff219a294a97a4ull,
    0x3ffffff847ab3b13ull, 0x4fffa470414a1ab1ull, 0x4fff03aab90a9784ull, 0x3ffffffab9a947a4ull,
    0x1ffffffffffff549ull, 0x2fffffffff830549ull, 0x2fffffffff501540ull, 0x3ffffff153358548ull,
    0x2fffffffff5492b1ull, 0x3ffffff9542b1083ull, 0x3ffffff0244252b5ull, 0x4fff483543253b52ull,
    0x2fffffffff3a2549ull, 0x3ffffff9548a0a20ull, 0x3ffffff3a2150540ull, 0x4fff8548a2582152ull,
    0x3ffffff54913b3abull, 0x4fffab8b18810954ull, 0x4fff03aab50a5405ull, 0x3ffffff8ab8b5485ull,
    0x2fffffffff795789ull, 0x3ffffff735539309ull, 0x3ffffff571170780ull, 0x2fffffffff573531ull,
    0x3ffffff12b579789ull, 0x4fff73530550912bull, 0x4fff52b578258028ull, 0x3ffffff573532b52ull,
    0x3ffffffa23897957ull, 0x4fff7a2209729579ull, 0x4fff5717811803a2ull, 0x3ffffff15717a21aull,
    0x4fff3ab13b578589ull, 0x5b0a0b1a07095705ull, 0x570507850b03ab0aull, 0x2fffffffffa57b5aull,
    0x1ffffffffffff65bull, 0x2fffffffffb65830ull, 0x2fffffffffb65019ull, 0x3ffffffb65981831ull,
    0x2fffffffff612651ull, 0x3ffffff083261651ull, 0x3ffffff260069659ull, 0x4fff283265825985ull,
    0x2fffffffff65b3a2ull, 0x3ffffff65b20a08aull, 0x3ffffffb653a2190ull, 0x4fff8a9a29921b65ull,
    0x3ffffff1355363a6ull, 0x4fffa65510a508a0ull, 0x4fff590650360a63ull, 0x3ffffff98a9a6596ull,
    0x2fffffffff784b65ull, 0x3ffffff5b6734304ull, 0x3ffffff478b65901ull, 0x4fff94773197165bull,
    0x3ffffff784516126ull, 0x4fff473043265251ull, 0x4fff260650059478ull, 0x5692965293947397ull,
    0x3ffffff65b847a23ull, 0x4fff7a2204724b65ull, 0x4fffb653a2784190ull, 0x5b65a474a9a29219ull,
    0x4fffa65513a53478ull, 0x54a0a470a1a651a5ull, 0x547863a360650590ull, 0x4fffa977949a6596ull,
    0x2fffffffff4b649bull, 0x3ffffff8309b4b64ull, 0x3ffffff40660b01bull, 0x4fff1b6648168318ull,
    0x3ffffff642241491ull, 0x4fff642492291083ull, 0x2fffffffff264240ull, 0x3ffffff264248328ull,
    0x3ffffff23a64b49bull, 0x4fffb649b48a2820ull, 0x4fff1b6640160a23ull, 0x5a181a28141b6416ull,
    0x4fff63a139369649ull, 0x541614961a108a18ull, 0x3ffffff640603a63ull, 0x2fffffffff68a486ull,
    0x3ffffff9b88b7b67ull, 0x4fff7b69b0b70730ull, 0x4fff801781b7167bull, 0x3ffffff73171b67bull,
    0x4fff678891681261ull, 0x5397930796912692ull, 0x3ffffff026067807ull, 0x2fffffffff726327ull,
    0x4fff67889b68b3a2ull, 0x5b797b69707a2072ull, 0x53a27b6b71781801ull, 0x4fff71661b17a21aull,
    0x536163a169678968ull, 0x2fffffffff67a910ull, 0x4fff60aa03067807ull, 0x1ffffffffffffa67ull,
    0x1ffffffffffff6a7ull, 0x2fffffffff76a083ull, 0x2fffffffff76a190ull, 0x3ffffff76a318198ull,
    0x2fffffffffa7612bull, 0x3ffffffa760832b1ull, 0x3ffffffa76b92902ull, 0x4fff98b83bb32a76ull,
    0x2fffffffff276237ull, 0x3ffffff206607087ull, 0x3ffffff190372762ull, 0x4fff768981861621ull,
    0x3ffffff37117b76bull, 0x4fff0818717b176bull, 0x4fffb76b907b0370ull, 0x3ffffffb98b876b7ull,
    0x2fffffffff86a846ull, 0x3ffffff4600636a3ull, 0x3ffffff0194686a8ull, 0x4fff36a319639469ull,
    0x3ffffffb12a86846ull, 0x4fff4606a00a32b1ull, 0x4fffb922906a4a84ull, 0x563436a43932b93bull,
    0x3ffffff624428238ull, 0x2fffffffff624420ull, 0x4fff384462342901ull, 0x3ffffff462421941ull,
    0x4fffb16468618138ull, 0x3ffffff04606b10bull, 0x593b390b36384634ull, 0x2fffffffffb4694bull,
    0x2fffffffff6a7954ull, 0x3ffffff76a954830ull, 0x3ffffff6a7405015ull, 0x4fff15354334876aull,
    0x3ffffff6a712b549ull, 0x4fff9548302b1a76ull, 0x4fff0242b44b56a7ull, 0x576a52b253543483ull,
    0x3ffffff495627237ull, 0x4fff876620860549ull, 0x4fff405501763623ull, 0x5581854182876286ull,
    0x4fff37176116b549ull, 0x55497080717616b1ull, 0x57b3b763b0b540b4ull, 0x4fff8b44b5b876b7ull,
    0x3ffffff89aa96956ull, 0x4fff9505606306a3ull, 0x4fff6a51505a0a80ull, 0x3ffffff315356a36ull,
    0x4fff56aa895a92b1ull, 0x52b16959606a0a30ull, 0x525052b05856a85aull, 0x4fff53bb32356a36ull,
    0x4fff823625285895ull, 0x3ffffff620609569ull, 0x5286823685801581ull, 0x2fffffffff162561ull,
    0x59686958636b1361ull, 0x4fff60550906b10bull, 0x2fffffffff6b5380ull, 0x1ffffffffffff56bull,
    0x2fffffffff5a75baull, 0x3ffffff30875a5baull, 0x3ffffff901ba5a75ull, 0x4fff318819a7b75bull,
    0x3ffffff51771a12aull, 0x4fff2a7751271830ull, 0x4fffa72029279759ull, 0x58292839252a7527ull,
    0x3ffffff7533525b2ull, 0x4fff25b758528208ull, 0x4fffb23375b35019ull, 0x552725b728219829ull,
    0x2fffffffff753351ull, 0x3ffffff751710870ull, 0x3ffffff375359039ull, 0x2fffffffff975879ull,
    0x3ffffffa8bb85845ull, 0x4fff30aba5a05045ull, 0x4fff45bba84b8190ull, 0x514341934a45ba4bull,
    0x4fff584a82852512ull, 0x51a5a125a4a304a0ull, 0x585a584a52590250ull, 0x2fffffffffa32459ull,
    0x4fff8434535235b2ull, 0x3ffffff204245b25ull, 0x51905848535b3b23ull, 0x4fff429921245b25ull,
    0x3ffffff513538458ull, 0x2fffffffff051450ull, 0x4fff350059538458ull, 0x1ffffffffffff459ull,
    0x3ffffffba99a4a74ull, 0x4fffba9a79974830ull, 0x4fff4a7401a41ba1ull, 0x5a4b4a7b41483143ull,
    0x4fff1292a9a49a74ull, 0x5830a121a9a79749ull, 0x3ffffff40242a74aull, 0x4fff24334842a74aull,
    0x4fff4973727929b2ull, 0x507270827b749b79ull, 0x50b4b014b7b237b3ull, 0x2fffffffff748b21ull,
    0x3ffffff137174914ull, 0x4fff718810174914ull, 0x2fffffffff437034ull, 0x1ffffffffffff874ull,
    0x2fffffffffa8bb89ull, 0x3ffffff9ba9a3093ull, 0x3ffffffba8b801b0ull, 0x2fffffffff3ba1b3ull,
    0x3ffffffa89a912a1ull, 0x4fffa922919a3093ull, 0x2fffffffff0a82a0ull, 0x1ffffffffffff2a3ull,
    0x3ffffff89b8b2382ull, 0x2fffffffff920b29ull, 0x4fffb811808b2382ull, 0x1ffffffffffffb21ull,
    0x2fffffffff189381ull, 0x1ffffffffffff910ull, 0x1ffffffffffff380ull, 0x0fffffffffffffffull,
};

static constexpr uint8_t kEdgeLo[12] = {0, 1, 3, 0, 4, 5, 7, 4, 0, 1, 3, 2};
static constexpr uint8_t kEdgeHi[12] = {1, 2, 2, 3, 5, 6, 6, 7, 4, 5, 7, 6};

/* Edge geometry relative to the cell (i,j,k): axis (0=x,1=y,2=z) and the offset of its lower
 * corner. Derived from kEdgeLo/kEdgeHi and the corner numbering above. */
static constexpr uint8_t kEdgeAxis[12] = {0, 1, 0, 1, 0, 1, 0, 1, 2, 2, 2, 2};
static constexpr uint8_t kEdgeDx[12]   = {0, 1, 0, 0, 0, 1, 0, 0, 0, 1, 0, 1};
static constexpr uint8_t kEdgeDy[12]   = {0, 0, 1, 0, 0, 0, 1, 0, 0, 0, 1, 1};
static constexpr uint8_t kEdgeDz[12]   = {0, 0, 0, 0, 1, 1, 1, 1, 0, 0, 0, 0};

MISO_HD constexpr int case_tri_count(uint64_t packed) { return (int)(packed >> 60); }

MISO_HD constexpr int case_edge(uint64_t packed, int slot) { return (int)((packed >> (4 * slot)) & 0xF); }

/* The edge geometry above as bit tests (no table load on the device). */
MISO_HD constexpr int edge_axis(int e) { return e >= 8 ? 2 : (e & 1); }
MISO_HD constexpr int edge_dx(int e) { return (0xA22 >> e) & 1; }
MISO_HD constexpr int edge_dy(int e) { return (0xC44 >> e) & 1; }
MISO_HD constexpr int edge_dz(int e) { return (0x0F0 >> e) & 1; }

}  // namespace miso

#endif
