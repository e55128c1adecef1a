// Device-resident acceleration structure shared by the builder and the
// traversal kernels (host and device compile this header).
//
// Layout in HBM (one scene):
//   nodes : QNode8[num_nodes], 80 B of node in each 128 B slot (one slot
//           per 128 B cache line, so a node visit never touches two lines),
//           breadth-first; the root is node 0 (node s for scene s of a
//           batched build); the internal children of a node are contiguous
//           from child_base.
//   tris  : TriRec[num_tris], 48 B each; the leaf children of a node are
//           contiguous from tri_base; a leaf ends at the record whose
//           TRI_LAST flag is set.
// A ray visit reads the 80 B of one QNode8 per node and one TriRec per
// triangle test, so the algorithmic bytes per ray are 48 (ray) + 36 (hit) +
// 80 n_node + 48 n_tri.
#pragma once
#include <stddef.h>
#include <stdint.h>

namespace rtg {

// 8-wide node with child boxes quantised to 8 bits per plane on a per-node
// power-of-two grid. Plane k of axis a decodes as fmaf(2^(e_a-127), q, p_a),
// and the builder checks with that exact expression that every decoded
// child box contains the child's widened float box, so the slab test is
// conservative. Slot c is empty when meta[c] == 0; otherwise it is an
// internal child (imask bit c, meta[c] = 0x80 | rank) at child_base + rank
// or a leaf (meta[c] = 1 + offset, at most 1 + 7 * 15) starting at
// tri_base + offset. Bit 7 of meta repeats imask so that a traversal can
// form a child's reference from its meta byte alone.
struct alignas(128) QNode8 {
  float px, py, pz;
  uint8_t ex, ey, ez;
  uint8_t imask;
  uint32_t child_base;
  uint32_t tri_base;
  uint8_t meta[8];
  uint8_t qlox[8], qhix[8];
  uint8_t qloy[8], qhiy[8];
  uint8_t qloz[8], qhiz[8];
  uint8_t pad[48];  // to the 128 B line
};
constexpr size_t QNODE_BYTES = 80;  // the part a visit reads
static_assert(sizeof(QNode8) == 128 && offsetof(QNode8, pad) == QNODE_BYTES, "QNode8: 80 B in a 128 B slot");


// Triangle record in Moeller-ready form, e1 = v0 - v1 and e2 = v2 - v0
// formed once in FP32 as the reference's leaf fill does
// (kernels/geometry/triangle.h:40-41,98-120).
//   w0 = (v0.x, v0.y, v0.z, primID bits)
//   w1 = (e1.x, e1.y, e1.z, geomID bits)
//   w2 = (e2.x, e2.y, e2.z, flags)
// Robust scenes keep the vertices themselves, as the reference's Triangle4v
// leaves do (kernels/geometry/trianglev.h): w1 = (v1, geomID), w2 = (v2, flags).
struct alignas(16) TriRec {
  float v0x, v0y, v0z;
  uint32_t primID;
  float e1x, e1y, e1z;
  uint32_t geomID;
  float e2x, e2y, e2z;
  uint32_t flags;
};
static_assert(sizeof(TriRec) == 48, "TriRec must be 48 bytes");

constexpr uint32_t TRI_LAST = 1u;

// Leaf record of a two-level scene's top tree: one instance of an object
// (a committed scene's own triangle tree). The traversal moves the ray into
// object space with m, as the reference's instance intersector does with
// the instance's world2local (kernels/geometry/instance_intersector.cpp:59-66,
// xfmPoint / xfmVector common/math/affinespace.h:102-103,
// common/math/linearspace3.h:156):
//   org' = fma(org.x, vx, fma(org.y, vy, fma(org.z, vz, p)))
//   dir' = fma(dir.x, vx, fma(dir.y, vy, dir.z * vz))
// per component; t keeps its meaning, so hits compare across instances.
struct alignas(16) InstRec {
  float m[12];      // world-to-object affine, columns vx | vy | vz | p
  uint32_t instID;  // instID[0] of hits below this record
  uint32_t object;  // index into the scene's object table
  uint32_t flags;   // TRI_LAST, INST_IDENTITY
  uint32_t pad;
};
static_assert(sizeof(InstRec) == 64, "InstRec must be 64 bytes");

// The record stands for the scene's own triangles: no transform, and hits
// report the query context's instID[0].
constexpr uint32_t INST_IDENTITY = 2u;

// Stack / child reference: bit 31 set = leaf (triangle record index).
constexpr uint32_t REF_LEAF = 0x80000000u;
// Binary-tree reference of a scene with no valid primitive (batched builds).
constexpr uint32_t REF_EMPTY = 0xfffffffeu;
// Stack entry that returns a two-level traversal from an object to the top tree.
constexpr uint32_t REF_INST_EXIT = 0xffffffffu;

// Per-node side record (same index as the QNode8): the float union box of
// the node's children and the number of primitives below it. Read by the
// learned split-cost trainer, not by the traversal.
struct NodeInfo {
  float lo[3];
  uint32_t count;
  float hi[3];
  uint32_t depth;
};
static_assert(sizeof(NodeInfo) == 32, "NodeInfo must be 32 bytes");

// One triangle geometry as the builder reads it (device pointers).
struct GeomDesc {
  const char* vertices;
  const char* indices;
  uint64_t vstride;
  uint64_t istride;
  uint32_t num_vertices;
  uint32_t num_tris;
  uint32_t prim_offset;  // first global primitive index of this geometry
  uint32_t geomID;
  uint32_t scene;        // scene of a batched build (0 otherwise)
  uint32_t kind;         // GEOM_TRIANGLES, or GEOM_INSTANCES: `vertices` holds one
                         // box (lo xyz, hi xyz) per primitive, `indices` its InstRec
};
constexpr uint32_t GEOM_TRIANGLES = 0u;
constexpr uint32_t GEOM_INSTANCES = 1u;

// What the traversal kernels need.
struct DevScene {
  const QNode8* nodes;
  const TriRec* tris;
  uint32_t num_nodes;
  uint32_t num_tris;
  float lo[3];     // scene box, for ray ordering keys
  float scale[3];  // 1023.99 / extent per axis (0 for a flat axis)
  uint32_t robust; // triangle records hold (v0, v1, v2) for the Pluecker test
  uint32_t root;   // root node of this scene (scenes of a batch share the arrays)
};

// What the two-level traversal needs: the top tree over instance records
// and the table of object trees they refer to.
struct DevTop {
  const QNode8* nodes;
  const InstRec* insts;
  const DevScene* objects;
  uint32_t num_nodes;
  uint32_t num_insts;
  uint32_t num_objects;
  uint32_t root;
  uint64_t max_iter;  // bound on the steps of one ray (corrupt-structure guard)
};

// Entries of the per-ray traversal stack; the builder rejects a tree whose
// deepest path could need more.
constexpr uint32_t TRACE_STACK = 512;

// Conservative-culling constants shared with the oracle's rules.
constexpr float CULL_MARGIN = 1.0f / 4096.0f;       // relative margin on ray interval ends
constexpr float BOX_EPS_REL = 1.0f / 1048576.0f;    // absolute box widening / max |coord|
constexpr float MIN_RCP_INPUT = 1e-18f;             // zero_fix threshold (reference constants.h:31)
constexpr float FLT_LARGE_ = 1.844e18f;             // valid vertex bound (reference isvalid)

}  // namespace rtg
