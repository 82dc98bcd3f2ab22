// Device-side descriptors of a verification batch: fragments, pairs, candidates, per-candidate
// accumulators and the Scores layout (crates/sherd-core/src/matching/verify.rs:73-160).
#pragma once

#include "common.hpp"
#include "grid.hpp"

namespace sherd {

// Scores as 20 doubles, in the reference's field order.
enum {
  VS_TIGHT_A, VS_TIGHT_B, VS_TIGHT, VS_GAP_A, VS_GAP_B, VS_GAP, VS_CONTACT_A, VS_CONTACT_B, VS_CONTACT, VS_SEAM,
  VS_GAP_LIMIT, VS_TIGHT_DELTA, VS_CONT, VS_CONT_N, VS_PEN, VS_PEN_DEPTH, VS_PEN_UNAVAILABLE, VS_PARTIAL, VS_BRK,
  VS_BRK_BEST, VS_COUNT
};

constexpr int MEDIAN_LDS = 2048;  // a median segment up to this many keys is selected from LDS (16 KB, so
                                  // the median's workgroups fit beside the side streams' kernels on a
                                  // CU; a larger segment is read from L2 on each pass)
constexpr uint32_t FLAG_VOXEL_RANGE = 1u;  // a seam voxel coordinate beyond +-2^20
constexpr uint32_t FLAG_GAVE_UP = 2u;      // a closest-approach query ran out of its step budget

// One fragment as R §6 reads it (verify.rs:223-255 `Surfaces`).
struct FragDesc {
  SceneView frac;
  SceneView mesh;
  int has_mesh;
  int watertight;
  double frac_area;
  const double *pf;
  const double *s;
  const double *brk_p;
  const double *brk_ns;
  const double *margin_p;
  const double *margin_n;
  // Per 64-point block of pf and of s (Morton order, so a block is a compact patch): the centre of
  // its box and the largest distance from it (x, y, z, r; f64, r rounded up). A wave's 64 lanes are
  // one block, so one sphere test can dismiss the whole wave's samples.
  const double *pf_blk;
  const double *s_blk;
  double pf_all[4];  // the bounding spheres of all of pf, s, brk_p and margin_p (x, y, z, r)
  double s_all[4];
  double brk_all[4];
  double margin_all[4];
  int n_pf, n_s, n_brk, n_margin;
  GridView margin;  // over margin_p, cell at least the largest `near` of the batch
  GridView brk;     // over brk_p, cell beyond the largest `seam` of the batch
};

struct PairDesc {
  int a, b;
  int pen_ok;     // both meshes present and watertight (verify.rs:490-493)
  double sc[12];  // Scales in scales.json order
};

struct CandDesc {
  int pair;
  int apart;           // pen_ok, and on both sides the whole sample set's sphere misses the other mesh's box
                       // (nothing can be inside: the closest approach is needed on both; set by k_prep)
  uint32_t seam_mask;  // seam voxel table capacity - 1
  size_t seam_off;     // into VerifyBatch::seam_table
  size_t seg[4];       // key segments: facing A, facing B, steps, agreements
  double T[16];
  double Ti[16];
};

struct CandAcc {
  uint32_t facing[2], tight[2], touching[2];
  uint32_t inside[2], deeper[2];
  uint32_t deepest[2];  // bits of the largest inside depth (non-negative f32)
  uint32_t closest[2];  // bits of the closest approach when nothing is inside
  uint32_t n_cont;
  uint32_t seam_cells;
  uint32_t flags;
  uint32_t in_box[2];  // samples that passed the mesh-box reject of the inside test, per side
  uint32_t pad;
  double med[4];
};

struct VerifyBatch {
  const FragDesc *frag;
  const PairDesc *pair;
  CandDesc *cand;
  CandAcc *acc;
  uint64_t *keys;        // median segments (f64 total-order keys)
  uint64_t *seam_table;  // per-candidate voxel hash sets
  double *scores;        // VS_COUNT per candidate
  uint8_t *accepted;
  const uint32_t *task;  // 2 * n_cand (candidate << 1 | side), grouped by the fragment queried
  uint64_t *hard;        // closest-approach queries left to whole waves: candidate << 33 | side << 32 | sample
  uint64_t *hard_a;      // the same for the apart candidates' pass
  uint32_t *hard_n;      // [0] queued, [1] taken, [2] inside samples listed in deep, [3] queued, [4] taken in hard_a,
                         // [5] groups listed in boxw
  uint32_t *boxq;        // per (candidate, side) task, box_cap entries from (2 candidate + side) box_cap: the
  int box_cap;           // samples inside the other mesh's box (CandAcc::in_box of them), whose rays k_pen_rays casts
  uint64_t *boxw;        // their groups of 64 entries: (2 candidate + side) << 32 | group
  uint32_t *gstart;      // per task (in task order) the first of its groups in boxw, and the total
  uint64_t *deep;        // inside samples whose depth k_pen_depth measures, as in hard
  float *deep_t;         // and their nearest ray crossings
  int n_cand;
  int full;
  int retest;
};

}  // namespace sherd
