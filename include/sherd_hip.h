/*
 * sherd_hip: the MI355X backend of sherd-refit's `Executor` boundary, as a C ABI.
 *
 * The reference's drop-in point is the Rust trait `Executor`
 * (crates/sherd-core/src/executor/mod.rs:105-137) with its batch structs
 * (crates/sherd-core/src/executor/batch.rs:42-371). Its implementations are CpuExecutor
 * (executor/cpu.rs) and the wgpu GpuExecutor (crates/sherd-gpu/src/executor.rs:219-487), selected
 * by sherd_backend::resolve (crates/sherd-backend/src/lib.rs:203-310). Every entry point below
 * names the reference interface it replaces.
 *
 * Conventions:
 *   - every function returns an int status: SHERD_OK (0) or a negative SHERD_E_* code; the
 *     message of the last failure on the calling thread is sherd_last_error();
 *   - all buffers are caller-owned host memory in the reference's own layouts: points and normals
 *     are n x 3 f64 row-major (`&[[f64; 3]]`), poses are 4 x 4 f64 row-major (`Matrix4<f64>`
 *     read row by row), meshes are f32 vertices and u32 faces (`RayScene` is f32);
 *   - nothing computes on the CPU: a call that cannot run on the device fails with a status, and
 *     the caller (the Rust shim) answers it with CpuExecutor, exactly as GpuExecutor delegates
 *     (crates/sherd-gpu/src/executor.rs:19-33);
 *   - a context serialises the calls made on it, so one context may be shared by threads (the
 *     trait is Send + Sync, executor/mod.rs:103-104); results never depend on the caller's
 *     thread or grouping.
 */
#ifndef SHERD_HIP_H
#define SHERD_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  SHERD_OK = 0,
  SHERD_E_ARG = -1,     /* a malformed argument (null pointer, size mismatch, bad enum) */
  SHERD_E_HIP = -2,     /* the HIP runtime reported an error */
  SHERD_E_NOMEM = -3,   /* a device allocation failed */
  SHERD_E_LIMIT = -4,   /* the input exceeds a documented device limit */
  SHERD_E_NODEV = -5,   /* no usable gfx950 device */
  SHERD_E_CORRUPT = -6, /* a result's receipt did not come back (the readback is not trusted) */
  SHERD_E_FORMAT = -7   /* a file that is missing, unreadable, or not a cache this build reads */
};

typedef struct sherd_ctx sherd_ctx;
typedef struct sherd_scene sherd_scene;

/* One point's answer to R §6.4 (batch.rs:342-352 `InsideOutcome`, 8 bytes). */
typedef struct {
  float depth;     /* distance to the surface when inside, 0 otherwise */
  uint32_t inside; /* 1 when inside */
} sherd_inside_outcome;

/* Per-method device counters (crates/sherd-gpu/src/executor.rs:50-354 `MethodStats`). */
typedef struct {
  uint64_t calls;      /* entry-point calls */
  uint64_t items;      /* points, rays, poses or candidates processed */
  uint64_t device_ns;  /* device time measured with HIP events around the kernels */
} sherd_method_stats;

enum {
  SHERD_M_FIRST_HIT, SHERD_M_BOUNDED_DISTANCE, SHERD_M_INSIDE, SHERD_M_COARSE, SHERD_M_ICP,
  SHERD_M_NEAREST, SHERD_M_VERIFY, SHERD_M_SCENE_BUILD, SHERD_M_CONE, SHERD_M_HUBER, SHERD_M_SEGMENT,
  SHERD_M_VERIFY_RAYS, SHERD_M_COUNT
};
/* SHERD_M_VERIFY_RAYS counts the parity rays of the inside tests that sherd_verify_batch runs (R
 * §6.4 through RayScene::inside, bvh.rs:146-158): items = 3 per surface sample whose moved position
 * passes the mesh-box reject, whichever two or three of its rays the majority vote actually needs;
 * calls = sherd_verify_batch calls with full = 1. Its rays run inside the verification kernels, so
 * device_ns stays 0 (their time is SHERD_M_VERIFY's). */

/* ---- the executor ------------------------------------------------------------------------- */

/* GpuExecutor::new + Gpu::open (crates/sherd-gpu/src/executor.rs:356, device.rs:528): binds the
 * device and creates the context's stream. */
int sherd_open(int device, sherd_ctx **out);
void sherd_close(sherd_ctx *ctx);
/* Executor::name (executor/mod.rs:107): "hip". */
const char *sherd_name(void);
const char *sherd_last_error(void);
/* The device marketing name and gfx target, for `info` lines (sherd-backend/src/lib.rs:30). */
int sherd_device_name(sherd_ctx *ctx, char *buf, size_t len);
/* Counters of one method (SHERD_M_*). */
int sherd_stats(sherd_ctx *ctx, int method, sherd_method_stats *out);

/* ---- the self-test gate and the failure contract --------------------------------------------- */

/* One row of the self-test (crates/sherd-gpu/src/selftest.rs:124-131 `Check`). */
typedef struct {
  char name[16];
  int passed;
  char detail[240];
} sherd_check;

/* What the self-test found (selftest.rs:135-150 `SelfTest`). */
typedef struct {
  int passed;                  /* every check held (selftest.rs:153 `passed`) */
  int n_checks;
  sherd_check checks[4];       /* limits, reduction, bounded nn, throughput, in that order */
  uint32_t reduce_device_bits; /* the f32 sum of the reduction check, device and host mirror */
  uint32_t reduce_host_bits;
  uint64_t queries;            /* bounded-NN queries the throughput figures are over */
  double gpu_ms, cpu_ms;       /* that batch: the kernel's host wall time, the host mirror's */
  double speedup;              /* cpu_ms / gpu_ms, what backend selection compares */
  int cpu_threads;
} sherd_selftest_report;

/* SelfTest::run (crates/sherd-gpu/src/selftest.rs:179), the gate of sherd_backend::resolve
 * (crates/sherd-backend/src/lib.rs:230-247): (1) limits: a gfx950 target, 64-lane waves, 64 KiB of
 * LDS per workgroup, 1024-lane workgroups; (2) a fixed-order f32 reduction of the 1e7 terms of
 * selftest.rs:405-418 over 256 workgroups of 256 lanes, then one, bit-identical to the host mirror
 * of the same order (selftest.rs:423-460; the reference's bits are 0x49a7230c); (3) the bounded
 * nearest-neighbour kernel of sherd_nearest_below on two 6000-point clouds on a warped unit sheet
 * under 64 small poses (r = 0.03): no hit/miss disagreement, |delta d| <= 1e-6, and at most 1e-5
 * of the queries choosing another neighbour, each a tie to 1e-6 (selftest.rs:83-120, 231-270);
 * (4) throughput: that batch's kernel wall time against the host mirror on every hardware thread
 * (recorded, never failed). A status other than SHERD_OK means the test could not run. */
int sherd_selftest(sherd_ctx *ctx, sherd_selftest_report *out);

/* The failure contract of crates/sherd-gpu/src/executor.rs:19-33, 425-487: every device failure is
 * per call and the shim answers that call from the CPU executor and counts it. These counters are
 * the library's side of that count (the shim keeps its own `delegated`). */
typedef struct {
  uint64_t calls;        /* entry-point calls on the context */
  uint64_t failed;       /* calls that returned a status other than SHERD_OK */
  uint64_t arg_errors;   /* SHERD_E_ARG */
  uint64_t host_errors;  /* SHERD_E_HIP */
  uint64_t nomem;        /* SHERD_E_NOMEM: a device allocation refused */
  uint64_t limit;        /* SHERD_E_LIMIT: a size policy or step budget */
  uint64_t corrupt;      /* SHERD_E_CORRUPT: a receipt that did not come back */
  uint64_t injected;     /* faults fired by sherd_fault_inject */
} sherd_health;
int sherd_health_read(sherd_ctx *ctx, sherd_health *out);

/* Fault injection for testing the shim's contract (the build's answer to executor.rs's host
 * error, corrupt readback and budget refusal triggers): arms one fault of `kind` to fire at its
 * `after`-th opportunity from now (0: the next), then disarms. SHERD_FAULT_HOST fires at an entry
 * point's prologue (the call returns SHERD_E_HIP having done nothing); SHERD_FAULT_NOMEM at a
 * request for device memory from the context's buffers (SHERD_E_NOMEM on the call's own error
 * path); SHERD_FAULT_CORRUPT at a receipt check (the receipt kernel writes the stale value and the
 * call returns SHERD_E_CORRUPT after its kernels ran). SHERD_FAULT_NONE disarms. Receipts are
 * checked by sherd_first_hit, sherd_bounded_distance, sherd_inside, sherd_nearest_below,
 * sherd_coarse_scores, the ICP entry points, sherd_verify_batch, sherd_match_stage1 and
 * sherd_huber_lm. The context keeps working after a fault. */
enum { SHERD_FAULT_NONE = 0, SHERD_FAULT_HOST = 1, SHERD_FAULT_NOMEM = 2, SHERD_FAULT_CORRUPT = 3 };
int sherd_fault_inject(sherd_ctx *ctx, int kind, int after);

/* ---- the fragment cache and MatchData (host side: what the pair stages read) ------------------ */

typedef struct sherd_cache sherd_cache;

/* The metadata block of a `.sherd` file (crates/sherd-core/src/fragment/cache.rs:104-176
 * `CacheMeta`) that the pair stages read, and the tensors' sizes. */
typedef struct {
  char name[256];
  uint32_t cache_version, target_faces, face_budget, n_boundary;
  int watertight;
  double thick, thick_mode, res, area, area0, frac_area;
  double brk_t, md_t;  /* brk_params.t and md_params.t: the t the arrays were built at */
  size_t nv, nf, n_brk, n_sub, n_s, n_pf, n_margin;
} sherd_cache_info;

/* cache::read / from_bytes (cache.rs:329-438, 384-389): a `.sherd` safetensors file whose
 * metadata says format "sherd-cache", cache_version 7 and algo_ref "2026-09-06/9d4b9d3"
 * (crates/sherd-core/src/lib.rs:130, 152), refused with SHERD_E_FORMAT and cache.rs's message
 * otherwise: a tensor of the wrong dtype or shape, a triangle outside V, labels that do not
 * describe F, breakline frames that do not describe brk_P, a brk_sub outside it, sp/fp that do not
 * describe S/Pf, a sample face outside F, a margin_idx outside S. Host memory only. */
int sherd_cache_read(const char *path, sherd_cache **out);
int sherd_cache_from_bytes(const uint8_t *bytes, size_t n, sherd_cache **out);
void sherd_cache_free(sherd_cache *cache);
int sherd_cache_info_get(const sherd_cache *cache, sherd_cache_info *out);

/* The stored tensors (cache.rs:11-26), copied into caller buffers sized by the info (any pointer
 * may be NULL): V nv x 3, F nf x 3, labels nf (0 shell, 1 fracture, 2 solid), brk_P / brk_ns /
 * brk_nf / brk_f n_brk x 3, brk_sub n_sub, brk_valid n_brk, S n_s x 3, sp n_s, Pf n_pf x 3, fp n_pf,
 * margin_idx n_margin. */
typedef struct {
  float *V;
  uint32_t *F;
  uint8_t *labels;
  float *brk_P, *brk_ns, *brk_nf, *brk_f;
  uint32_t *brk_sub;
  uint8_t *brk_valid;
  float *S;
  uint32_t *sp;
  float *Pf;
  uint32_t *fp;
  uint32_t *margin_idx;
} sherd_cache_arrays_out;
int sherd_cache_arrays(const sherd_cache *cache, const sherd_cache_arrays_out *out);

/* Sizes of MatchData::own's arrays. */
typedef struct {
  size_t n_reg, n_reg_frac, n_reg_margin; /* pc_reg and its two prefixes */
  size_t n_pf, n_brk, n_sub, n_margin, n_s, nf;
  double t, frac_area;
  int watertight;
} sherd_match_info;

/* MatchData::own's arrays (each pointer may be NULL), f32 values widened to f64 exactly:
 * pc_reg (reg_p, reg_n: n_reg x 3), pc_frac (pf, pf_n = FN[fp]: n_pf x 3), the breakline (brk_p,
 * brk_ns, brk_f: n_brk x 3; brk_t = ns x f in f64, unnarrowed; brk_dih in degrees: n_brk), the
 * hypothesis subset (brk_sub: n_sub), the margin (margin_p = S[margin_idx], margin_n =
 * FN[sp[margin_idx]]: n_margin x 3), the surface samples (s, s_n = FN[sp]: n_s x 3; s_frac =
 * labels[sp] == fracture: n_s), and the face normals FN (nf x 3). */
typedef struct {
  double *reg_p, *reg_n;
  double *pf, *pf_n;
  double *brk_p, *brk_ns, *brk_f, *brk_t, *brk_dih;
  uint32_t *brk_sub;
  double *margin_p, *margin_n;
  double *s, *s_n;
  uint8_t *s_frac;
  double *face_normals;
} sherd_match_arrays;

/* MatchData::own (crates/sherd-core/src/fragment/samples.rs:499-560) at the fragment's own t: FN
 * from face_geometry in f64 over the f32 vertices, narrowed (types.rs:100-112, geometry.rs:70-96);
 * pc_reg the prefixes of registration_split(n_pf, n_margin, reg_points) (samples.rs:585-615;
 * numpy's half-to-even rounding, at least one fracture point; reg_points 0 keeps all, the
 * reference's default is 6000). With out NULL only info is filled. SHERD_E_LIMIT when the cached
 * arrays were built at another t than the fragment's (the reference rebuilds those,
 * samples.rs:562-575). */
int sherd_match_data(const sherd_cache *cache, size_t reg_points, sherd_match_info *info,
                     const sherd_match_arrays *out);

/* ---- scenes (RayScene, crates/sherd-core/src/spatial/bvh.rs:36-82) -------------------------- */

/* RayScene::of_mesh / of_subset: a device LBVH over f32 vertices V (nv x 3) and faces F (nf x 3).
 * mask (nf bytes, nonzero keeps the face) may be NULL; a mask that keeps nothing falls back to the
 * first face (bvh.rs:71-76). Face indices reported by queries are positions in the kept list. */
int sherd_scene_build(sherd_ctx *ctx, const float *V, size_t nv, const uint32_t *F, size_t nf,
                      const uint8_t *mask, sherd_scene **out);
void sherd_scene_free(sherd_scene *scene);
/* RayScene::n_faces and RayScene::aabb (bvh.rs:182-194). */
int sherd_scene_info(const sherd_scene *scene, size_t *faces, float lo[3], float hi[3]);

/* RayScene::first_hit (bvh.rs:84-91): n rays (origins and directions n x 3 f32); smallest t with
 * ties to the lowest face; a miss is t = +inf, prim = 0xFFFFFFFF. retest != 0 retests
 * near-degenerate hits in FP64. */
int sherd_first_hit(sherd_ctx *ctx, const sherd_scene *scene, const float *orig, const float *dir,
                    size_t n, int retest, float *t, uint32_t *prim);

/* Executor::bounded_distance (executor/mod.rs:121; CpuExecutor at cpu.rs:84-113; DistBatch at
 * batch.rs:285-310): points moved by T in f64, narrowed to f32, distance to the scene when at most
 * max_dist. reduce = 0 (DistReduce::All): n values, +inf beyond the window. reduce = 1
 * (DistReduce::Min): one value, the minimum of the window and every distance. */
int sherd_bounded_distance(sherd_ctx *ctx, const sherd_scene *scene, const double *pts, size_t n,
                           const double T[16], double max_dist, int reduce, double *out);

/* No counterpart in the reference. sherd_bounded_distance without any allocation, for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the same kernels and results, with every buffer beyond the caller's host
 * arrays and the scene in a fixed 9 MiB device array per device (taken at sherd_open). At most
 * 262144 points per call (SHERD_E_LIMIT above). Calls on one device share that array: calls that
 * overlap on one device (two contexts on the same GPU, from two threads) race; calls through one
 * context are serialised like every other entry point. */
int sherd_bounded_distance_noalloc(sherd_ctx *ctx, const sherd_scene *scene, const double *pts, size_t n,
                                   const double T[16], double max_dist, int reduce, double *out);

/* Executor::inside (executor/mod.rs:124; cpu.rs:115-128, 182-188; InsideBatch at batch.rs:316):
 * the AABB reject, three axis rays with a majority of odd crossing counts, the depth of an inside
 * point. retest = 0 is the reference's arithmetic; 1 adds the FP64 retest of degenerate hits. */
int sherd_inside(sherd_ctx *ctx, const sherd_scene *scene, const double *pts, size_t n,
                 const double T[16], int retest, sherd_inside_outcome *out);

/* ---- nearest neighbours (PointTree, crates/sherd-core/src/spatial/kdtree.rs) ---------------- */

/* PointTree::nearest_below (kdtree.rs:182-192) over an f64 cloud (n x 3): for each of m queries
 * the nearest point with ties to the lowest index, when sqrt(d^2) < bound; idx = -1 otherwise.
 * d^2 is kiddo's ((dx*dx + dy*dy) + dz*dz) in f64. */
int sherd_nearest_below(sherd_ctx *ctx, const double *cloud, size_t n, const double *q, size_t m,
                        double bound, int64_t *idx, double *dist);

/* ---- the coarse score (Executor::coarse_scores) --------------------------------------------- */

/* Executor::coarse_scores (executor/mod.rs:115; cpu.rs:46-68, 136-169; CoarseBatch at
 * batch.rs:148-176): poses are k x 12 f64, the rotation row-major then the translation (the
 * `Poses::get` view, batch.rs:110-126, of both Split and Homogeneous poses). counts (may be NULL)
 * gets the agree count per pose; scores gets f64(count) / n_probe, and 0 when the probe is empty. */
int sherd_coarse_scores(sherd_ctx *ctx, const double *target_p, const double *target_n,
                        size_t n_target, const double *probe_p, const double *probe_n,
                        size_t n_probe, const double *poses, size_t k, double radius,
                        double normal_agree, uint32_t *counts, double *scores);

/* ---- one ICP rung (Executor::icp_rung) -------------------------------------------------------- */

/* `Registration` (crates/sherd-core/src/matching/icp.rs:182-196). */
typedef struct {
  double transform[16]; /* row-major 4x4 */
  double fitness;
  double inlier_rmse;
  uint64_t correspondences;
  uint32_t iterations;
  uint32_t converged;
} sherd_registration;

enum { SHERD_POINT_TO_POINT = 0, SHERD_POINT_TO_PLANE = 1 };

/* Executor::icp_rung (executor/mod.rs:118; cpu.rs:70-82 -> icp::register, icp.rs:308-356;
 * IcpBatch at batch.rs:188-238): k candidates from `inits` (k x 16), each up to max_iteration
 * iterations of Open3D's ICP at `distance`, in f64. tgt_n may be NULL (a point-to-plane rung then
 * makes identity updates, icp.rs:460-462). */
int sherd_icp_rung(sherd_ctx *ctx, const double *src, size_t n_src, const double *tgt_p,
                   const double *tgt_n, size_t n_tgt, const double *inits, size_t k,
                   int estimation, double distance, int max_iteration, sherd_registration *out);

/* ---- batched ICP: rungs and ladders over many cloud pairs (ladder.rs, icp.rs) ------------------ */

typedef struct sherd_cloud sherd_cloud;

/* IcpTarget::new (crates/sherd-core/src/matching/icp.rs:225-240) and the moving clouds of a rung
 * (ladder.rs:69-80 `Rung::source`): n x 3 f64 points and, optionally, unit normals (NULL for a
 * source or a point-to-point target), copied to the device with a point LBVH that stands in for
 * the KD-tree. At most 262144 points. */
int sherd_cloud_create(sherd_ctx *ctx, const double *pts, const double *normals, size_t n, sherd_cloud **out);
void sherd_cloud_free(sherd_cloud *cloud);

/* One cloud pair of a batch and its correspondence radius (`Scales::icp_dist(k)` of the pair,
 * scales.rs:100, ladder.rs:116). */
typedef struct {
  const sherd_cloud *source;
  const sherd_cloud *target;
  double distance;
} sherd_icp_pair;

/* A rung (ladder.rs:69 `Rung`): the update and the iteration cap. */
typedef struct {
  int estimation; /* SHERD_POINT_TO_POINT or SHERD_POINT_TO_PLANE */
  int max_iteration;
} sherd_icp_rung_spec;

/* ladder::climb_all (crates/sherd-core/src/matching/ladder.rs:105-128) for many pairs at once:
 * candidate i starts from inits[16 i ..] on pair cand_pair[i] and climbs every rung, each rung from
 * the previous rung's transform. pairs holds n_rungs x n_pairs entries, rung-major (a rung may use
 * other clouds, as R §5.6's pc_reg and pc_frac rungs do). out receives n_cand x n_rungs
 * registrations, candidate-major. Each rung is one persistent launch; the rungs are captured once as
 * a graph and replayed while the batch's shape is unchanged. Results do not depend on the batch's
 * composition or the workgroup size: each candidate's rung equals sherd_icp_rung on its own. */
int sherd_icp_batch_ladder(sherd_ctx *ctx, const sherd_icp_pair *pairs, size_t n_pairs,
                           const sherd_icp_rung_spec *rungs, size_t n_rungs, const uint32_t *cand_pair,
                           const double *inits, size_t n_cand, sherd_registration *out);
/* One rung of the above (Executor::icp_rung over many pairs). */
int sherd_icp_batch_rung(sherd_ctx *ctx, const sherd_icp_pair *pairs, size_t n_pairs, const uint32_t *cand_pair,
                         const double *inits, size_t n_cand, int estimation, int max_iteration,
                         sherd_registration *out);
/* Fixes the form of every ICP rung: the persistent batched kernel at 64, 128, 256 or 512 threads
 * per workgroup (one candidate per workgroup at a time), or 1, the wide form (each pass a grid over
 * all spans of all candidates, one thread per source point: the form for a few candidates on large
 * clouds). With 0 the autotuner times every form on the first rung of each (source-size,
 * candidate-count) bucket and keeps the fastest. Every form gives the same registrations to the
 * bit. No counterpart in the reference (its wgpu kernels have fixed tiles). */
int sherd_icp_block(sherd_ctx *ctx, int threads);

/* ---- R §5.1-5.5 over many pairs: hypotheses, suppression, the stage-1 hand-off ----------------- */

typedef struct sherd_frames sherd_frames;

/* hypotheses::Frames (crates/sherd-core/src/matching/hypotheses.rs:47-68): one fragment's breakline
 * frames (n points: brk_P, shell normals ns, in-plane axes f, tangents t, all f64 n x 3; dihedral
 * angles dih in degrees) and its hypothesis subset brk_sub (n_sub indices, in the order the caller
 * means; PMC-4), on the device with a neighbour grid over the points for the coarse scores. */
int sherd_frames_create(sherd_ctx *ctx, const double *p, const double *ns, const double *f, const double *t,
                        const double *dih, size_t n, const uint32_t *sub, size_t n_sub, sherd_frames **out);
void sherd_frames_free(sherd_frames *frames);

/* hypotheses::build_with (hypotheses.rs:149-165) and poses (:172-212): the frame pairs of the
 * subsets ia (of A) and ib (of B) whose dihedral angles satisfy |dih_A + dih_B - 180| < tol,
 * row-major over (ia, ib); pa / pb are positions into ia / ib; poses get R (row-major) then tau,
 * 12 f64 each, in the reference's summation order. *n gets the count; fails if it exceeds
 * capacity. */
int sherd_hypotheses(sherd_ctx *ctx, const sherd_frames *a, const sherd_frames *b, const uint32_t *ia, size_t nia,
                     const uint32_t *ib, size_t nib, double dihedral_tol, size_t capacity, uint32_t *pa, uint32_t *pb,
                     double *poses, size_t *n);

/* nms::nms (crates/sherd-core/src/matching/nms.rs:62-84): the greedy walk over `order` (item
 * indices) of items with poses (R row-major then tau, 12 f64) and scores: stop at the first score
 * under floor; drop an item nearer than trans_tol to a kept one and turned by trace(R^T R') > 2.9;
 * stop once topk are kept (tested after each item, so topk = 0 keeps one). kept receives the kept
 * items in the order kept (room for min(n_order, topk + 1)). */
int sherd_nms(sherd_ctx *ctx, const uint32_t *order, size_t n_order, const double *poses, const double *score,
              size_t n_items, double trans_tol, size_t topk, double floor, uint32_t *kept, size_t *n_kept);

/* One pair of a matching batch: A fixed, B moving (Pair, crates/sherd-core/src/matching/pair.rs). */
typedef struct {
  const sherd_frames *a;
  const sherd_frames *b;
  const uint32_t *ia;        /* the subsets the hypotheses pair, or NULL for the frames' own */
  size_t nia;
  const uint32_t *ib;
  size_t nib;
  const double *probe_p;     /* R §5.2's probe of B (Probe::draw, pair.rs:175-177): n_probe x 3 */
  const double *probe_n;
  size_t n_probe;
  const sherd_cloud *s1_source; /* R §5.4's ladder: B's pc_brk, A's pc_brk_full (with a point tree) */
  const sherd_cloud *s1_target;
  double scales[12];         /* Scales in scales.json order */
} sherd_match_pair;

typedef struct {
  uint32_t n_hyp;            /* hypotheses of the pair */
  uint32_t n_kept1;          /* poses R §5.3 kept (stage 1) */
  uint32_t n_kept2;          /* stage-1 poses R §5.5 kept (stage 2) */
  uint32_t pad;
  double best1;              /* the best stage-1 score (0 without a stage-1 pose) */
} sherd_match_result;

/* Pair::suppress, stage1 and suppress_stage1 (pair.rs:188-264; hypotheses.rs:139-141;
 * coarse::scores; ladder.rs:131-139, 177-210) for n_pairs pairs in the same launches: hypotheses,
 * coarse scores at Scales::coarse, the walk (descending score, ties by index; PMC-6), suppression
 * (5000 walked, stage1 kept, floor 0.1), the two point-to-point rungs (0.2, 0.08; 20 iterations)
 * from every kept hypothesis, the re-score of B's subset at Scales::stage1, and the second
 * suppression (stage2 kept, floor 0.05). Per pair p: res[p]; kept1[p*stage1 + k] the kept
 * hypotheses (indices into the pair's hypothesis list); s1_T[16 (p*stage1 + k)] and
 * s1_score[p*stage1 + k] their stage-1 transforms and scores; kept2[p*stage2 + k] the stage-2
 * poses as positions into the pair's kept1 list. dihedral_tol: Params::dihedral_tol (25). */
int sherd_match_stage1(sherd_ctx *ctx, const sherd_match_pair *pairs, size_t n_pairs, size_t stage1, size_t stage2,
                       double dihedral_tol, sherd_match_result *res, uint32_t *kept1, double *s1_T, double *s1_score,
                       uint32_t *kept2);

/* ---- R §9 full-resolution refinement (crates/sherd-core/src/refine.rs) ------------------------- */

/* refine.rs:216-227 fracture_cloud: the cloud of an original mesh (V f64 nv x 3, F u32 nf x 3) whose
 * vertices have a fracture face of the working mesh as nearest face centroid (centroids f64
 * n_working_faces x 3, `face_geometry`'s (a + b + c) / 3, geometry.rs:88-93; fracture: one byte per
 * working face) strictly closer than max(0.15 thick, 1.5 res) (select_candidates, refine.rs:163-182),
 * ascending, then capped at max_points by a seeded draw without replacement (cap_selection,
 * refine.rs:186-192; rng.rs:127-156); normals are Open3D's area-weighted vertex normals
 * (vertex_normals, refine.rs:118-150). The cloud has no point tree (place it with one to use it as
 * an ICP target). */
int sherd_fracture_cloud(sherd_ctx *ctx, const double *V, size_t nv, const uint32_t *F, size_t nf,
                         const double *centroids, const uint8_t *fracture, size_t n_working_faces, double thick,
                         double res, size_t max_points, uint64_t seed, sherd_cloud **out);
/* refine.rs:207-214 cloud_from_indices with the mesh's vertex normals: the vertices idx, in order. */
int sherd_cloud_from_indices(sherd_ctx *ctx, const double *V, size_t nv, const uint32_t *F, size_t nf,
                             const uint32_t *idx, size_t n, sherd_cloud **out);
/* The number of points of a cloud, and a copy of its vertex indices (fracture clouds only; NULL to
 * skip), points and normals (NULL to skip). */
size_t sherd_cloud_size(const sherd_cloud *cloud);
int sherd_cloud_read(const sherd_cloud *cloud, uint32_t *idx, double *points, double *normals);
/* FractureCloud::placed (refine.rs:99-104): a new cloud, points through the 4x4 T (row-major) and
 * normals through its rotation, each row the fused chain of types.rs:202-214; with_tree builds the
 * point LBVH an ICP target needs. */
int sherd_cloud_placed(const sherd_cloud *cloud, const double T[16], int with_tree, sherd_cloud **out);

/* RefinePiece (refine.rs:237-245): the fragment's wall and resolution, and its cloud (NULL: none). */
typedef struct {
  double thick;
  double res;
  const sherd_cloud *cloud;
} sherd_refine_piece;

/* RefinedJoin (refine.rs:248-262). */
typedef struct {
  uint32_t fixed;
  uint32_t moving;
  double dist[2];
  double rungs[2][16];
  double fitness;
  double rmse_t;
} sherd_refined_join;

/* refine.rs:270-303 refine_joins: per group (members flattened, sizes per group) of two or more,
 * the walk over `used` (n_used (a, b) pairs in assembly order) from the group's first fragment;
 * each join's moving cloud placed at its pose and the fixed cloud at its final pose, registered
 * by two point-to-plane rungs at Scales::icp_dist(0.05) and (0.02), 40 iterations each, from the
 * identity (refine_one, refine.rs:306-350); poses_out[moving] = rung 2 . poses[moving]. The joins
 * of one depth of the walk run as one device ladder. joins (capacity: one per used join) receives
 * the joins taken in walk order, *n_joins their count; a join whose either side has no cloud is
 * skipped and leaves its pose alone. */
int sherd_refine_joins(sherd_ctx *ctx, const sherd_refine_piece *pieces, size_t n_pieces, const double *poses,
                       const uint32_t *members, const size_t *sizes, size_t n_groups, const uint32_t *used,
                       size_t n_used, double *poses_out, sherd_refined_join *joins, size_t *n_joins);

/* ---- the robust seam energy and its se(3) gradient (new) --------------------------------------- */

/* New: no reference function computes this energy; its nearest relative is ICP's point-to-plane
 * normal equations (crates/sherd-core/src/matching/icp.rs:599-640). For pose i (base transform T0,
 * twist xi = (omega, v), rotation first as icp.rs:432-433 orders x) and frozen correspondences
 * s_j -> (q, n):  p_j = exp(omega^) T0 s_j + v,  r_j = (p_j - q) . n,  E = sum rho_delta(r_j)
 * (Huber), dE/dv = sum psi(r_j) n, dE/domega = J_l(omega)^T sum psi(r_j) (exp(omega^) T0 s_j) x n.
 * At xi = 0 and delta -> infinity the gradient is normal_equations' JTr (icp.rs:629). */
typedef struct sherd_huber sherd_huber;

/* A batch of k poses: pose i on pair cand_pair[i] (the pair's target must carry normals; its
 * distance is the correspondence radius) from T0[16 i ..]. The correspondences are taken at T0
 * (icp.rs:361-392) and frozen. The clouds must outlive the handle. */
int sherd_huber_create(sherd_ctx *ctx, const sherd_icp_pair *pairs, size_t n_pairs, const uint32_t *cand_pair,
                       const double *T0, size_t k, sherd_huber **out);
void sherd_huber_free(sherd_huber *h);
/* The frozen correspondences, target index or -1, poses in order (n = the sum of their source sizes). */
int sherd_huber_correspondences(sherd_huber *h, int32_t *corr, size_t n);
/* E (k) and, if grad is not NULL, dE/dxi (k x 6) at the twists xi (k x 6). precision 32 computes the
 * per-point terms and sums in f32 about the clouds' centroids; 64 in f64. */
int sherd_huber_eval(sherd_huber *h, const double *xi, double delta, int precision, double *energy, double *grad);
/* Batched Levenberg-Marquardt on E from T0: each iteration re-takes the correspondences at the
 * current pose, solves the reweighted 6x6 system in f64 and keeps the step if E drops (the drop
 * summed point by point, sum rho(r') - rho(r), which f32 still resolves near the minimum). T_out k x 16,
 * energy k (at the final pose, over its last correspondences), steps k (accepted iterations). */
int sherd_huber_lm(sherd_huber *h, double delta, int iterations, int precision, double *T_out, double *energy,
                   uint32_t *steps);

/* ---- R §6 verification (crates/sherd-core/src/matching/verify.rs) ------------------------------ */

typedef struct sherd_fragment sherd_fragment;

/* One fragment as R §6 measures it (verify.rs:223-302 `Surfaces`): its fracture scene and whole
 * mesh scene (borrowed handles; mesh may be NULL, which makes the penetration test unavailable),
 * R §3.3.2's watertight verdict, R §3.4's fracture area, and its sample sets, all n x 3 f64:
 * pf (fracture samples), s (surface samples), brk_p / brk_ns (breakline and shell normals),
 * margin_p / margin_n (shell margin and normals). The arrays are copied to the device. */
int sherd_fragment_create(sherd_ctx *ctx, const sherd_scene *fracture, const sherd_scene *mesh,
                          int watertight, double frac_area, const double *pf, size_t n_pf,
                          const double *s, size_t n_s, const double *brk_p, const double *brk_ns,
                          size_t n_brk, const double *margin_p, const double *margin_n,
                          size_t n_margin, sherd_fragment **out);
void sherd_fragment_free(sherd_fragment *frag);

/* A pair of fragments and its Scales (scales.rs:29-54, scales.json order: t, res, coarse, stage1,
 * tight, facing, gap, seam, near, pen, nms, icp). */
typedef struct {
  const sherd_fragment *a;
  const sherd_fragment *b;
  double scales[12];
} sherd_verify_pair;

enum {
  SHERD_S_TIGHT_A, SHERD_S_TIGHT_B, SHERD_S_TIGHT, SHERD_S_GAP_A, SHERD_S_GAP_B, SHERD_S_GAP,
  SHERD_S_CONTACT_A, SHERD_S_CONTACT_B, SHERD_S_CONTACT, SHERD_S_SEAM, SHERD_S_GAP_LIMIT,
  SHERD_S_TIGHT_DELTA, SHERD_S_CONT, SHERD_S_CONT_N, SHERD_S_PEN, SHERD_S_PEN_DEPTH,
  SHERD_S_PEN_UNAVAILABLE, SHERD_S_PARTIAL, SHERD_S_BRK, SHERD_S_BRK_BEST, SHERD_S_COUNT
};

/* verify::verify (verify.rs:590-626) and verify::accept (verify.rs:632-638, Params::default) for
 * n_cand candidates over many pairs in one batch: candidate i is pose cand_T[16 i ..] of pair
 * cand_pair[i]. scores: n_cand x SHERD_S_COUNT doubles (the Scores fields in order; brk and
 * brk_best are left 0 for the caller); accepted: n_cand bytes. full = 0 is R §5.6's cheap half.
 * retest = 1 adds the FP64 retest of degenerate parity hits. This replaces the per-candidate
 * `par_iter` over verify in Pair::stage2_batch (crates/sherd-core/src/matching/pair.rs:350-359). */
int sherd_verify_batch(sherd_ctx *ctx, const sherd_verify_pair *pairs, size_t n_pairs,
                       const uint32_t *cand_pair, const double *cand_T, size_t n_cand, int full,
                       int retest, double *scores, uint8_t *accepted);

/* ---- preprocessing rays (R §3.4.3; the trait's fifth method, executor/mod.rs:9-11) ------------ */

/* segment::classify_faces (crates/sherd-core/src/fragment/segment.rs:318-381): the seven-ray cone
 * vote of each listed face. scene is the mesh's own LBVH (built from V narrowed to f32, no mask);
 * V (nv x 3 f64) and F (nf x 3) are the same mesh; faces lists n_faces face indices (NULL: all
 * faces, n_faces = nf); ns holds one f64 shell normal per listed face; thick is the wall t.
 * good receives the vote count 0..7 per listed face. */
int sherd_cone_votes(sherd_ctx *ctx, const sherd_scene *scene, const double *V, size_t nv,
                     const uint32_t *F, size_t nf, const int64_t *faces, const double *ns,
                     size_t n_faces, double thick, int retest, uint8_t *good);

/* The per-vertex form of the preprocessing rays (the benchmark's slab plate casts one ray per vertex
 * along the inward normal; the reference's own form is the cone above, segment.rs:318-381, whose
 * first ray and test this one applies per vertex): vertex v's normal is Open3D's area-weighted
 * vertex normal of V (nv x 3 f64) and F (refine.rs:118-150), its ray starts 1e-3 inside along it
 * (narrowed to f32) and goes straight down; good[v] = 1 when the first hit lies in the wall window
 * (0.1 t < dh, 0.5 < dh / t < 1.8 in f32) on a face that looks back along the ray (h . d > 0.7 in
 * f64). frac (nf bytes, may be NULL): 1 for a face fewer than two of whose vertices voted. scene is
 * the mesh's own LBVH (V narrowed to f32, no mask). */
int sherd_vertex_votes(sherd_ctx *ctx, const sherd_scene *scene, const double *V, size_t nv, const uint32_t *F,
                       size_t nf, double thick, int retest, uint8_t *good, uint8_t *frac);

/* ---- R §3.2, the wall thickness ---------------------------------------------------------------- */

/* thickness::estimate_thickness (crates/sherd-core/src/fragment/thickness.rs:100-110, with
 * cast_inward_rays :115-132 and thickness_from_hits :139-161): one inward ray from the centroid of
 * every face of R §3.2's stride (all of them up to 300 000, else every ceil(nf / 300000)-th), the
 * f64 face normals of V (nv x 3) and F (nf x 3), and the histogram modes of the hits. scene is the
 * mesh's own LBVH (V narrowed to f32, no mask). On return *estimated is 1 with *t (the filtered
 * mode) and *thick_mode (the mode over every valid hit), or 0 when fewer than 100 rays hit (the
 * reference's None, whose caller falls back to the oriented box). */
int sherd_thickness(sherd_ctx *ctx, const sherd_scene *scene, const double *V, size_t nv, const uint32_t *F,
                    size_t nf, float *t, float *thick_mode, int *estimated);
/* thickness::thickness_from_hits (thickness.rs:139-161) on given rays: n_rays hits (t_hit f32,
 * +inf on a miss; prim, 0xFFFFFFFF on a miss) of the rays cast from faces idx, over a mesh of nf
 * faces with f64 unit normals (nf x 3). Outputs as sherd_thickness. */
int sherd_thickness_from_hits(sherd_ctx *ctx, const double *normals, size_t nf, const uint32_t *idx,
                              const float *t_hit, const uint32_t *prim, size_t n_rays, float *t,
                              float *thick_mode, int *estimated);

/* ---- R §3.4.4-3.4.8, the clean-up of the cone vote's mask -------------------------------------- */

/* Optional outputs of sherd_segment_cleanup (each pointer may be null): the mask after the
 * majority filter and after the two island passes (nf bytes each), the fixed shell reference
 * normal and whether its ball held a shell face (nf x 3 f64, nf bytes), and the number of growth
 * passes that found candidates. The fields of segment.rs:146-171 SegTrace this step produces. */
typedef struct {
  uint8_t *frac_majority;
  uint8_t *frac_islands;
  double *reference;
  uint8_t *has_ref;
  int passes;
} sherd_seg_trace;

/* segment.rs:544-635, steps 3.4.4 to 3.4.8 of segment::run after the cone vote, on the working
 * mesh V (nv x 3 f64), F (nf x 3) with wall t = thick: the majority filter (majority_filter,
 * segment.rs:403-428; balls of t/4 around the voxel representatives rep[n_rep], a face taking its
 * nearest representative near[f]), fracture islands below 0.5 t^2 and shell islands below 2 t^2
 * (mesh/components.rs:170-205 over mesh/adjacency.rs:138-158's pairs), the shell's reference
 * normal over balls of t/2 (ball_normals, segment.rs:267-305), the boundary growth within
 * boundary_angle degrees (refine_boundary, segment.rs:438-480; the default is 25), and the last
 * fracture island pass. frac_raw is the vote's mask (1 = fracture); labels receives the final mask.
 * rep and near are the voxel grid of coarse_grid (segment.rs:189-205). A component's area is the
 * exact sum of its faces' areas in units of 2^(E-64), 2^(E-1) <= t^2 < 2^E; the reference sums in
 * face order, which decides differently only for a component within its rounding of the bound.
 * A non-positive or non-finite thick labels every face shell, as segment.rs:520-529 does. */
int sherd_segment_cleanup(sherd_ctx *ctx, const double *V, size_t nv, const uint32_t *F, size_t nf, double thick,
                          const uint8_t *frac_raw, const uint32_t *rep, size_t n_rep, const uint32_t *near,
                          double boundary_angle, uint8_t *labels, sherd_seg_trace *trace);

/* R §3.4.2's smoothed normals (segment.rs:540-546 with ball_normals, :267-305): per voxel
 * representative rep[q] the area-weighted mean normal of every face whose centroid lies within
 * radius of its centroid (kiddo's squared distance, the ball summed in ascending face order),
 * normalised with the reference's two 1e-12 floors; face f receives its representative near[f]'s
 * (ns, nf x 3 f64). The reference's radius is max(t/3, smooth_res res). */
int sherd_smoothed_normals(sherd_ctx *ctx, const double *V, size_t nv, const uint32_t *F, size_t nf,
                           const uint32_t *rep, size_t n_rep, const uint32_t *near, double radius, double *ns);

/* ---- per-kernel timing ------------------------------------------------------------------------ */

/* Turns per-kernel HIP-event timing on the context's stream on (1) or off (0), clearing the
 * totals. Counterpart of the reference's per-method device counters (sherd-gpu/src/executor.rs). */
int sherd_profile(sherd_ctx *ctx, int on);
/* The totals as a JSON object {"kernel": {"ms": total, "launches": n}, ...}. */
int sherd_profile_read(sherd_ctx *ctx, char *buf, size_t len);


#ifdef __cplusplus
}
#endif

#endif
