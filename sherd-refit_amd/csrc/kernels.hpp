// Host-side declarations of the device launchers (one translation unit per subsystem).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "common.hpp"
#include "grid.hpp"
#include "verify.hpp"

namespace sherd {

// Optional per-kernel timing: HIP events recorded on the launching stream around each kernel,
// read back once the stream has synchronised.
struct KernelProfiler {
  bool on = false;
  static constexpr int MAXK = 64;
  const char *name[MAXK] = {};
  double ms[MAXK] = {};
  uint64_t launches[MAXK] = {};
  int n = 0;
  hipEvent_t ev[2 * MAXK] = {};
  int pending[MAXK] = {};
  int npending = 0;
  int slot(const char *k);
  void begin(const char *k, hipStream_t s);  // record the start event of kernel k
  void end(hipStream_t s);                    // record the end event of the last begun kernel
  void collect();                             // after a stream synchronisation
};

// A 4x4 pose passed by value to a kernel (row-major f64).
struct Pose {
  double m[16];
};

// ---- LBVH (bvh.hip) ----------------------------------------------------------------------------
struct DevScene {
  Node *nodes = nullptr;
  Tri *tris = nullptr;
  AxisGrid pg[3] = {};  // built on demand by scene_parity_grids
  DistBound db = {};    // built on demand by scene_distance_bounds
  int n = 0;
  int root = -1;
  float lo[3] = {0, 0, 0};
  float hi[3] = {0, 0, 0};
  float absmax = 0;
};

// Builds the tree over m device faces (dF, 3 u32 each, indexing dV's f32 triples). Synchronises
// the stream before returning (the root and the bounds come back to the host).
hipError_t scene_build(const float *dV, const uint32_t *dF, int m, DevScene *out, hipStream_t s);
void scene_free(DevScene *s);
SceneView view_of(const DevScene &d);
// Builds the scene's three parity grids unless built; synchronises the stream.
hipError_t scene_parity_grids(DevScene *sc, hipStream_t s);
// Builds the scene's distance bounds unless built; synchronises the stream.
hipError_t scene_distance_bounds(DevScene *sc, hipStream_t s);

hipError_t launch_first_hit(const SceneView &sv, const float *orig, const float *dir, int n, int retest, float *t,
                            uint32_t *prim, hipStream_t s);
hipError_t launch_inside(const SceneView &sv, const double *pts, int n, const Pose &T, int retest, float *depth,
                         uint32_t *inside, hipStream_t s);
hipError_t launch_bounded(const SceneView &sv, const double *pts, int n, const Pose &T, float window, float *out,
                          hipStream_t s);
hipError_t launch_bounded_min(const SceneView &sv, const double *pts, int n, const Pose &T, unsigned int *best,
                              hipStream_t s);

// ---- f64 neighbour grid (grid.hip) ------------------------------------------------------------
struct DevGrid {
  double *pts = nullptr;
  uint32_t *idx = nullptr;
  GridSlot *table = nullptr;
  uint32_t mask = 0;
  int n = 0;
  double origin[3] = {0, 0, 0};  // the points' box: origin is its low corner
  double hi[3] = {0, 0, 0};
  double cell = 1.0;
};

// Builds the grid over n device points (3 f64 each) for queries of reach <= `reach`.
// Synchronises the stream (the bounding box comes back to size the cells).
hipError_t grid_build(const double *dp, int n, double reach, DevGrid *g, hipStream_t s);
void grid_free(DevGrid *g);
GridView view_of(const DevGrid &d);

hipError_t launch_nearest(const GridView &g, const double *q, int m, double bound, int64_t *idx, double *dist,
                          hipStream_t s);
hipError_t launch_coarse(const GridView &g, const double *tn, const double *pp, const double *pn, int np,
                         const double *poses, int k, double radius, double agree, uint32_t *counts, double *scores,
                         hipStream_t s);

// ---- ICP rung (icp.hip) --------------------------------------------------------------------------
// One workgroup per candidate runs every iteration of the rung (icp::register, icp.rs:308-356).
// moved: k * ns * 3 f64 scratch; corr: k * ns u32 scratch; out: k * 21 doubles.
// ---- batched ICP (icp.hip) ----------------------------------------------------------------------
constexpr int ICP_NV = 29;        // per-pass sums: count, err^2, then 27 update sums
constexpr int ICP_SPAN = 512;     // source points per span of the fixed summation order
constexpr int ICP_MAX_SPANS = 512;  // so at most 262144 source points per cloud

// One cloud pair of a batch: the moving source (points only) and the target with its point LBVH
// and (for point-to-plane) normals, and the pair's correspondence radius.
struct IcpPairDesc {
  const double *src;
  int ns;
  const double *tp;
  const double *tn;
  SceneView tree;
  double r;
  double cs[3], ct[3];  // the source's and the target's centroids
};

struct IcpBatchView {
  const IcpPairDesc *pairs;
  const uint32_t *cand_pair;
  const uint32_t *order;  // candidates, largest source first
  const size_t *off;      // each candidate's first point in moved / corr
  const double *T_in;     // 16 per candidate
  double *T_out;          // 16 per candidate (may be null)
  double *out;            // 21 per candidate: transform, fitness, rmse, correspondences, iterations, converged
  double *moved;
  uint32_t *corr;
  double *part;           // per workgroup: part_spans x ICP_NV span sums
  int part_spans;
  uint32_t *next;         // queue position
  int n_cand;
  int est;                // 0 point-to-point, 1 point-to-plane, 2 identity updates
  int max_iter;
};

// ---- Huber energy and LM (huber.hip) ----------------------------------------------------------------
// One pose's LM state between the launches of an iteration.
struct LmState {
  double T[16];
  double lambda, E, Ef;
  int done, steps;
};

struct HuberView {
  const IcpPairDesc *pairs;
  const uint32_t *cand_pair;
  const size_t *off;     // each pose's first source point in corr / lm_corr
  const double *T0;      // 16 per pose
  const int32_t *corr;   // frozen correspondences (target index or -1)
  int32_t *lm_corr;      // the LM's own correspondences
  LmState *lm;           // per pose
  int k;
  int max_ns;            // the largest source cloud of the batch
};

hipError_t launch_correspond(const HuberView &v, const double *Tc, int32_t *corr, hipStream_t s);
hipError_t launch_huber_eval(const HuberView &v, const double *xi, double delta, int f64, double *E, double *G,
                             hipStream_t s);
hipError_t launch_huber_lm(const HuberView &v, double delta, int iters, int f64, double *T_out, double *E_out,
                           uint32_t *steps, hipStream_t s);

// Workgroup sizes the batched kernel is built for (the autotuner's choices); returns their count.
int icp_block_sizes(int *out);
// Resets the queue and launches one persistent rung over the batch.
hipError_t launch_icp_batch(const IcpBatchView &v, int block, int grid, hipStream_t s);

// The wide form of a rung (a few candidates, large clouds): every span of every candidate is one
// workgroup of a pass. Per candidate its registration state between the launches of a rung.
struct IcpWideState {
  double T[16], U[16];
  double red[ICP_NV];
  double fit, rmse;
  int stop, iters, conv, pad;
};
struct IcpWideView {
  IcpBatchView b;               // b.order lists the candidates that run, b.n_cand of them
  const uint32_t *span_first;   // per position in b.order its first span, then the total (n_cand + 1)
  int n_spans;
  IcpWideState *st;             // per candidate
  double *part;                 // per span: ICP_NV sums
  int k0, s0;                   // the first queue position and span of a launch (its grid's offset)
};
// How a rung's queue splits into chains: n parts, part j from queue position k[j] and span s[j]
// (k[0] = s[0] = 0; k[n], s[n] the ends). One part when the rung is too small to split.
constexpr int ICP_CHAINS = 2;
struct IcpWideSplit {
  int n = 1;
  int k[ICP_CHAINS + 1] = {}, s[ICP_CHAINS + 1] = {};
};
// Queues the rung's max_iteration + 1 passes; results land in b.out / b.T_out as for the batched form.
// With a split, the parts of the queue run their passes as chains on s and on side[0..n-2] (forked by
// `fork`, joined by join[0..n-2]), so one part's correspondence pass overlaps another's update;
// every candidate's arithmetic is the same.
hipError_t launch_icp_wide(const IcpWideView &w, hipStream_t s, KernelProfiler *prof, const IcpWideSplit &split,
                           const hipStream_t *side, hipEvent_t fork, const hipEvent_t *join);

// ---- R §5.1-5.5 over many pairs (stage.hip) -----------------------------------------------------
// One fragment's breakline frames on the device (hypotheses.rs:47-68 `Frames`).
struct FramesView {
  const double *p, *ns, *f, *t, *dih;
  const double *sub_p, *sub_n;  // the subset's points and shell normals (R §5.4's re-score probe)
  int n, n_sub;
  GridView grid;                // over p, cell at least every radius it is queried at
};
// One pair of a matching batch: A is the fixed fragment (its breakline the target), B the moving.
struct StagePair {
  FramesView a, b;
  const uint32_t *ia, *ib;      // the subsets the hypotheses pair (their positions are pa / pb)
  int nia, nib;
  const double *probe_p, *probe_n;  // R §5.2's probe of B
  int n_probe;
  double dihedral_tol, coarse_r, stage1_r;
};
// One greedy suppression (nms.rs:62-84) over items of a global pose / score table.
struct NmsProblem {
  const uint32_t *order;  // the walk, item indices
  int n;
  const double *pose;     // 12 per item: R row-major, then tau
  const double *score;
  double tol;
  int topk;
  double floor;
  uint32_t *kept;         // item indices, in the order kept
  uint32_t *n_kept;
};
hipError_t launch_hyp_count(const StagePair *sp, const uint32_t *row_first, int n_pairs, uint32_t rows, uint32_t *cnt,
                            hipStream_t s);
hipError_t launch_hyp_build(const StagePair *sp, const uint32_t *row_first, int n_pairs, uint32_t rows,
                            const uint32_t *off, uint32_t n_hyp, uint32_t *pa, uint32_t *pb, uint32_t *hpair,
                            double *pose, hipStream_t s);
hipError_t launch_stage_coarse(const StagePair *sp, const uint32_t *ipair, const double *pose, uint32_t n, int which,
                               uint32_t *counts, double *scores, hipStream_t s);
// Sorts items by (pair, count descending, index within the pair ascending); vals gets item indices.
hipError_t launch_order(const uint32_t *ipair, const uint32_t *item_first, const uint32_t *counts, uint32_t n,
                        uint64_t *keys, uint32_t *vals, uint64_t *tk, uint32_t *tv, uint32_t *hist, hipStream_t s);
size_t order_hist_words(uint32_t n);
hipError_t launch_nms(const NmsProblem *probs, int n, hipStream_t s);
hipError_t launch_pose12(const double *T, uint32_t n, double *pose, hipStream_t s);
hipError_t launch_gather12(const double *pose, const uint32_t *idx, uint32_t n, double *out, hipStream_t s);

// ---- R §9 refinement (refine.hip) ---------------------------------------------------------------
// Vertex selection of an original mesh (nv f64 vertices) against the working mesh's face centroids
// C (point LBVH `tree`) and fracture labels: sel/pos per vertex, pos scanned exclusive.
hipError_t launch_select(const double *V, int nv, const SceneView &tree, const double *C, const uint8_t *fracture,
                         double radius, uint8_t *sel, uint32_t *pos, uint32_t *scan_part, hipStream_t s);
size_t select_scan_words(int nv);
hipError_t launch_compact(const uint8_t *sel, const uint32_t *pos, int nv, uint32_t *idx, hipStream_t s);
// Points and vertex normals of the n vertices idx of a mesh. fn: 3 nf f64 scratch; scratch:
// normals_scratch_words(nv, nf) u32.
size_t normals_scratch_words(int nv, int nf);
hipError_t launch_vertex_cloud(const double *V, int nv, const uint32_t *F, int nf, const uint32_t *idx, int n,
                               double *fn, uint32_t *scratch, double *pts, double *nrm, hipStream_t s);
// A cloud through a pose: points by the 4x4, normals (when both pointers are set) by its rotation.
hipError_t launch_place(const double *p, const double *n, int cnt, const Pose &T, double *po, double *no,
                        hipStream_t s);

// ---- verification (verify.hip) -------------------------------------------------------------------
// Three more streams (default priority) and the events that fork them from the main stream and join
// them back: the independent sections of a verification run side by side.
struct SideStreams {
  hipStream_t side[3] = {};
  hipEvent_t fork = nullptr;
  hipEvent_t early = nullptr;  // the fork of the apart candidates' closest approach, after k_prep
  hipEvent_t join[3] = {};
};

// max_pf / max_s / max_margin: the largest per-fragment sample counts in the batch (grid sizes).
// With `ss`, R §6.1 (k_frac), R §6.2 (k_seam) and R §6.3 (k_cont) run on its side streams while
// R §6.4 runs on s; without, everything runs on s in turn.
hipError_t launch_verify(const VerifyBatch &vb, int max_pf, int max_s, int max_margin, int max_brk, hipStream_t s,
                         const SideStreams *ss, KernelProfiler *prof);

// ---- cone vote (cone.hip) -------------------------------------------------------------------------
struct ConeConst {
  double cos_a, sin_a;  // of the 15 degree half-angle
  double cp[6], sp[6];  // of the six cone azimuths 2 pi k / 6
  float min_hit;        // (0.1 t) as f32
  float thick;          // t as f32
};

// The per-vertex rays (cone.hip): good[v] = 1 when vertex v's inward ray meets the opposite wall;
// frac (may be null) per face, 1 when fewer than two of its vertices did.
hipError_t launch_vertex_votes(const SceneView &sv, const double *V, int nv, const uint32_t *F, int nf,
                               const double *nrm, double thick, int retest, uint8_t *good, uint8_t *frac,
                               hipStream_t s);
hipError_t launch_iota(uint32_t *x, int n, hipStream_t s);
hipError_t launch_cone(const SceneView &sv, const double *V, const uint32_t *F, int nf, const int64_t *faces,
                       const double *ns, int n, const ConeConst &cc, int retest, uint8_t *good, double *fn,
                       hipStream_t s);

// ---- R §3.4.4-3.4.8 (segment.hip) -----------------------------------------------------------------
constexpr int SEG_BALL_CAP = 1024;  // faces within one representative's ball
hipError_t launch_seg_geom(const double *V, int nv, const uint32_t *F, int nf, double *N, double *A, double *C,
                           uint32_t *bad, hipStream_t s);
hipError_t launch_seg_edge_keys(const uint32_t *F, int nf, uint32_t nv, uint64_t *key, uint32_t *pos, hipStream_t s);
hipError_t launch_seg_pairs(const uint64_t *key, const uint32_t *pos, int m, int nf, uint32_t *fa, uint32_t *fb,
                            uint32_t *count, hipStream_t s);
// err bits: 1 a representative outside the faces, 2 a ball over SEG_BALL_CAP faces, 4 a face's
// representative outside the list.
hipError_t launch_seg_majority(const GridView &g, const double *C, const double *A, const uint8_t *frac_raw, int nf,
                               const uint32_t *rep, int n_rep, const uint32_t *near, double r, uint8_t *maj,
                               uint8_t *frac, uint32_t *err, hipStream_t s);
hipError_t launch_seg_reference(const GridView &g, const double *C, const double *A, const double *N,
                                const uint8_t *frac, int nf, const uint32_t *rep, int n_rep, const uint32_t *near,
                                double r, double cos_grow, double *ref_g, uint8_t *has_g, double *ref, uint8_t *has,
                                uint8_t *elig, uint32_t *err, hipStream_t s);
// R §3.4.2: the area-weighted normal of every face within r of each representative, per face.
hipError_t launch_seg_smoothed(const GridView &g, const double *C, const double *A, const double *N, int nf,
                               const uint32_t *rep, int n_rep, const uint32_t *near, double r, double *ref_g,
                               uint8_t *has_g, double *ns, uint32_t *err, hipStream_t s);
hipError_t launch_cc_init(int n, uint32_t *parent, hipStream_t s);
hipError_t launch_cc_round(const uint32_t *fa, const uint32_t *fb, int np, const uint8_t *mask, uint8_t target, int n,
                           uint32_t *parent, uint32_t *changed, hipStream_t s);
hipError_t launch_cc_drop(int n, uint8_t *mask, uint8_t target, const uint32_t *parent, const double *A, int sh,
                          unsigned long long *acc, uint64_t bound_hi, uint64_t bound_lo, hipStream_t s);
hipError_t launch_grow_pass(const uint32_t *fa, const uint32_t *fb, int np, int n, uint8_t *frac, uint8_t *cand,
                            const uint8_t *elig, uint32_t *flags, hipStream_t s);

// ---- R §3.2 (thick.hip) -----------------------------------------------------------------------------
// R §3.2's device state, zeroed before each estimate: the two list counts, the two modes, a flag
// raised by a face index outside V, and per list the selection's and the histogram's counts.
constexpr int THICK_BINS = 60;
struct ThickSel {
  uint32_t hist[4][256];   // per digit pass, the keys with the selected prefix by their next digit
  uint32_t gt_inv;         // ~(smallest key above the lower order statistic), 0 if none
  uint32_t bins[THICK_BINS];
  float first, last;       // the histogram's range
  uint32_t pad;
};
struct ThickState {
  uint32_t counts[2];  // valid hits, looking-back hits
  float modes[2];      // their histogram modes (NaN when empty)
  uint32_t bad_face;
  uint32_t pad[3];
  ThickSel sel[2];
};
hipError_t launch_thick_normals(const double *V, int nv, const uint32_t *F, int nf, double *N, uint32_t *bad,
                                hipStream_t s);
hipError_t launch_thick_rays(const SceneView &sv, const double *V, int nv, const uint32_t *F, const double *N,
                             int stride, int m, float *t_hit, uint32_t *prim, hipStream_t s);
// The valid and the looking-back distances into ok_d / far_d (m floats each), their counts and
// histogram modes into st (zeroed by the caller).
hipError_t launch_thick_modes(const double *N, int nf, const uint32_t *idx, int stride, const float *t_hit,
                              const uint32_t *prim, int m, float *ok_d, float *far_d, ThickState *st, hipStream_t s);

}  // namespace sherd
