/*
 * The oracle: a CPU restatement of sherd-refit's CpuExecutor and the R §6 verification, used as
 * TEST INFRASTRUCTURE ONLY.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load this library, and only
 * as the checker (or as the timed CPU baseline). Nothing in the product links it.
 *
 * Every function cites the reference file:line it restates (paths relative to the reference
 * repository root). Two parts have no bit-level counterpart the reference can pin, because the
 * reference delegates them to un-vendored crates (parry3d 0.30.2, kiddo 6.2.0, nalgebra 0.34.2):
 *
 *   - the closest-point-on-triangle and first-hit arithmetic (parry's `project_local_point*`,
 *     `cast_local_ray`): this restatement defines its own f32 algorithms, pinned by closed forms,
 *     brute force and the slab fixture at the reference's tolerances ("parity unpinned" at the bit
 *     level against parry);
 *   - the 3x3 SVD inside Umeyama (nalgebra `svd`): a one-sided Jacobi SVD, pinned by exact rigid
 *     recoveries and the slab stage-1 poses at tolerance.
 *
 * Everything else (Moller-Trumbore parity, the kNN semantics, coarse counts, the ICP loop, the
 * LDLT with Eigen pivots, pose_inverse, median, scales, the verification statistics and the
 * accept rule) is restated operation for operation in the reference's order.
 */
#ifndef SHERD_ORACLE_H
#define SHERD_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_scene oracle_scene;

/* RayScene::of_mesh / of_subset (crates/sherd-core/src/spatial/bvh.rs:56-82).
 * V: nv x 3 f32, F: nf x 3 u32, mask: nf bytes or NULL. An empty mask selection falls back to
 * face 0 (bvh.rs:71-76). Returns NULL for an empty face list. */
oracle_scene *oracle_scene_new(const float *V, size_t nv, const uint32_t *F, size_t nf,
                               const uint8_t *mask);
void oracle_scene_free(oracle_scene *s);
size_t oracle_scene_faces(const oracle_scene *s);
/* RayScene::aabb (bvh.rs:182-186): min/max over the scene's vertices. */
void oracle_scene_aabb(const oracle_scene *s, float lo[3], float hi[3]);

/* First hit (bvh.rs:84-91 contract): smallest t, ties to the lowest scene-local face.
 * Miss: t = +inf, prim = 0xFFFFFFFF. retest != 0 enables the FP64 retest of near-degenerate
 * hits. brute != 0 scans every triangle instead of the BVH (used to pin the BVH). */
void oracle_first_hit(const oracle_scene *s, const float *orig, const float *dir, size_t n,
                      int retest, int brute, float *t_out, uint32_t *prim_out);
/* RayScene::crossings (bvh.rs:160-178): Moller-Trumbore hits along one ray, all of them. */
uint32_t oracle_crossings(const oracle_scene *s, const float o[3], const float d[3], int retest,
                          int brute);
/* Unsigned distance, unbounded (bvh.rs:110-113). */
float oracle_distance(const oracle_scene *s, const float p[3], int brute);

/* CpuExecutor::bounded_distance (crates/sherd-core/src/executor/cpu.rs:84-113):
 * points n x 3 f64 moved by T (4x4 row-major f64) and narrowed to f32. reduce_min = 0 writes n
 * values (+inf beyond max_dist); reduce_min = 1 writes one value. */
void oracle_exec_bounded_distance(const oracle_scene *s, const double *pts, size_t n,
                                  const double T[16], double max_dist, int reduce_min,
                                  double *out);
/* CpuExecutor::inside (cpu.rs:115-128, 182-188): depth f32 and inside u32 per point. */
void oracle_exec_inside(const oracle_scene *s, const double *pts, size_t n, const double T[16],
                        int retest, float *depth, uint32_t *inside);

/* PointTree::nearest_below (crates/sherd-core/src/spatial/kdtree.rs:182-192) over an f64 cloud:
 * idx = -1 on a miss. */
void oracle_nearest_below(const double *cloud, size_t n, const double *q, size_t m, double bound,
                          int64_t *idx, double *dist);
/* PointTree::nearest_distance (crates/sherd-core/src/spatial/kdtree.rs:120-123, kiddo's
 * nearest_one): the unbounded nearest of each query by brute force, squared distance
 * ((dx*dx + dy*dy) + dz*dz), ties to the lowest index; dist = sqrt of it. n must be > 0. */
void oracle_nearest_unbounded(const double *cloud, size_t n, const double *q, size_t m, int64_t *idx,
                              double *dist);
/* apply_transform_fused / rotate_fused (crates/sherd-core/src/types.rs:202-214): each row
 * fma(t2, p2, fma(t1, p1, t0 * p0)) (+ t3 for points, not for directions). */
void oracle_place(const double *p, size_t n, const double T[16], int direction, double *out);
/* The f32 mirror of HashGrid::nearest_below / nearest_within (spatial/grid.rs:600-650),
 * brute force: strict or inclusive d^2 against r*r in f32, ties to the lowest index. */
void oracle_nearest_f32(const float *cloud, size_t n, const float *q, size_t m, float radius,
                        int strict, int64_t *idx, float *dist);

/* CpuExecutor::coarse_scores (cpu.rs:46-68, 136-169). poses: k x 12 f64, row-major R (9) then tau
 * (3) -- the `Poses::get` view (executor/batch.rs:110-126). counts (may be NULL) gets the agree
 * count per pose; scores gets f64(count)/n. */
void oracle_coarse_scores(const double *target_p, const double *target_n, size_t n_target,
                          const double *probe_p, const double *probe_n, size_t n_probe,
                          const double *poses, size_t k, double radius, double normal_agree,
                          uint32_t *counts, double *scores);

/* icp::register over a batch (crates/sherd-core/src/matching/icp.rs:308-356; cpu.rs:70-82).
 * estimation: 0 point-to-point, 1 point-to-plane. inits: k x 16 row-major.
 * out: k x 21 doubles: T(16), fitness, rmse, correspondences, iterations, converged. */
void oracle_icp_rung(const double *src, size_t n_src, const double *tgt_p, const double *tgt_n,
                     size_t n_tgt, const double *inits, size_t k, int estimation, double distance,
                     int max_iteration, double *out);

/* icp.rs:361-392 correspondences of already-moved source points: corr[i] = the nearest target
 * strictly inside `distance`, or -1. */
void oracle_icp_correspondences(const double *moved, size_t n, const double *tgt_p, size_t n_tgt, double distance,
                                int64_t *corr);
/* icp.rs:599-640 normal_equations (world assembly) over the pairs of `corr`: JTJ (6x6, row-major)
 * and JTr for rows [p x n, n] and residuals (p - q) . n, summed in source order. */
void oracle_normal_equations(const double *moved, size_t n, const double *tgt_p, const double *tgt_n,
                             const int64_t *corr, double jtj[36], double jtr[6]);

/* One fragment as R §6 measures it (crates/sherd-core/src/matching/verify.rs:223-302). */
typedef struct {
  const oracle_scene *fracture;
  const oracle_scene *mesh; /* NULL: penetration unavailable */
  int watertight;
  double frac_area;
  const double *pf;
  size_t n_pf;
  const double *s;
  size_t n_s;
  const double *brk_p;
  const double *brk_ns;
  size_t n_brk;
  const double *margin_p;
  const double *margin_n;
  size_t n_margin;
} oracle_side;

/* The Scores struct (verify.rs:73-160) as 20 doubles, in this order. */
enum {
  OS_TIGHT_A, OS_TIGHT_B, OS_TIGHT, OS_GAP_A, OS_GAP_B, OS_GAP, OS_CONTACT_A, OS_CONTACT_B,
  OS_CONTACT, OS_SEAM, OS_GAP_LIMIT, OS_TIGHT_DELTA, OS_CONT, OS_CONT_N, OS_PEN, OS_PEN_DEPTH,
  OS_PEN_UNAVAILABLE, OS_PARTIAL, OS_BRK, OS_BRK_BEST, OS_COUNT
};

/* Scales in scales.json order: t, res, coarse, stage1, tight, facing, gap, seam, near, pen, nms,
 * icp (crates/sherd-core/src/matching/scales.rs:29-54). */
void oracle_scales_for_pair(double t, double res, double out[12]);

/* The self-test's fixed-order f32 reduction (crates/sherd-gpu/src/selftest.rs:405-460): the bits
 * of the sum of reduce_terms(n) over `groups` workgroups, then one. */
uint32_t oracle_reduce_bits(size_t n, size_t groups);
/* verify::verify (verify.rs:590-626). retest: FP64 retest in the parity rays. */
void oracle_verify(const oracle_side *a, const oracle_side *b, const double T[16],
                   const double scales[12], int full, int retest, double scores[OS_COUNT]);
/* verify::accept (verify.rs:632-638) under Params::default (params.rs:165-169). */
int oracle_accept(const double scores[OS_COUNT], const double scales[12]);
/* verify::pose_inverse (verify.rs:739-798). */
void oracle_pose_inverse(const double T[16], double out[16]);
/* verify::median (verify.rs:693-700); sorts `values` in place. */
double oracle_median(double *values, size_t n);

/* segment::classify_faces (crates/sherd-core/src/fragment/segment.rs:318-381) over `faces`
 * (indices into F), with the f64 face geometry of V (mesh/geometry.rs:70-96). ns: one f64 normal
 * per listed face. good: one vote count per listed face. */
void oracle_cone_votes(const oracle_scene *s, const double *V, size_t nv, const uint32_t *F,
                       size_t nf, const int64_t *faces, const double *ns, size_t n_faces,
                       double thick, int retest, uint8_t *good);

/* The batched Huber LM (oracle/huber_lm.cpp) for k poses of one pair: source n_src x 3, target
 * n_tgt x 3 with unit normals, radius, base poses T0 (k x 16). T_out (k x 16), E_out, steps (last
 * accepted iteration), status (0 cap, 1 converged, 2 damped out, 3 no step). threads: 0 = all. */
void oracle_huber_lm(const double *src, size_t n_src, const double *tp, const double *tn, size_t n_tgt,
                     double radius, const double *T0, size_t k, double delta, int iterations, int threads,
                     double *T_out, double *E_out, uint32_t *steps, int32_t *status);

/* Threads the parallel loops use (rayon's axes); 0 = all. Returns the count in force. */
int oracle_set_threads(int n);

#ifdef __cplusplus
}
#endif

#endif
