// Device BVH node records of libblazert_amd (shared by the engine's kernels and by device code that traverses a
// tree built through brt_bvh_build, such as the user-collection traversal of include/blazert_amd/user_geometry.h).
//
// The tree is the reference's binary binned-SAH tree (blazert/bvh/builder.h:84-115, numbered in pre-order) stored
// "children in parent": an internal node record holds its two child boxes, so one record fetch drives both slab
// tests of a traversal step; the root box is kept beside the records. A child reference is an internal-record index
// or LEAF_BIT | first primitive (a position in the leaf order), with the leaf's primitive count beside it.
// Plain C++ structs: usable from host code and HIP device code alike.
#pragma once
#include <stdint.h>

namespace brt {

constexpr uint32_t LEAF_BIT = 0x80000000u;
constexpr uint32_t AXIS_SHIFT = 30;           // cnt[1] high bits carry the parent's split axis
constexpr uint32_t CNT_MASK = (1u << AXIS_SHIFT) - 1u;
constexpr uint32_t INVALID_ID = 0xFFFFFFFFu;

// FP64 internal node: 128 B, one cache line. Child c box: b[6c+0..2] = min xyz, b[6c+3..5] = max xyz.
struct alignas(128) Node64 {
  double b[12];
  uint32_t ref[2];  // child reference: internal node index, or LEAF_BIT | first primitive (leaf order)
  uint32_t cnt[2];  // leaf primitive count (0 for an internal child); cnt[1] >> 30 = split axis
  uint32_t pad[4];
};
static_assert(sizeof(Node64) == 128, "Node64 must be one 128-byte line");

// The same record with FP32 boxes (64 B) for trees built in T = float (blazert::Scene<float>, BVH<float, ...>).
struct alignas(64) Node32 {
  float b[12];
  uint32_t ref[2];
  uint32_t cnt[2];
};
static_assert(sizeof(Node32) == 64, "Node32 size");

}  // namespace brt
