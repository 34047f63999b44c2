/*
 * angelo_amd -- C-ABI of the MI355X (gfx950) NeuS / Neuralangelo hot path.
 *
 * Conventions (all entry points):
 *   - return int status: AM_OK (0) or a negative error class; am_last_error()
 *     gives the thread-local message of the last failure.  No exception ever
 *     crosses this boundary.
 *   - every device buffer is allocated by the caller (PyTorch's caching
 *     allocator) and passed as a raw pointer plus sizes; the library never
 *     allocates device memory.  Scratch comes from a caller workspace whose
 *     size the matching *_workspace_size query returns.
 *   - launches are asynchronous on the caller's HIP stream (`stream`, a
 *     hipStream_t passed as void*); nothing synchronises the device.
 *   - arguments are validated on the host before any launch.
 *   - fp32 everywhere unless the name says otherwise; row-major arrays.
 *
 * Reference interfaces replaced are cited as path:line into hugoycj/
 * Instant-angelo (tiny-cuda-nn and nerfacc 0.3.3 are not vendored there;
 * their call sites are cited).
 */
#ifndef ANGELO_AMD_H
#define ANGELO_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define AM_OK 0
#define AM_ERR_INVALID -1     /* bad argument / shape / size */
#define AM_ERR_HIP -2         /* HIP runtime error */
#define AM_ERR_WORKSPACE -3   /* workspace too small */
#define AM_ERR_UNSUPPORTED -4 /* configuration not implemented */
#define AM_ERR_INTERNAL -5    /* internal consistency check failed */

#define AM_MAX_LEVELS 32

const char* am_last_error(void);
/* After AM_ERR_WORKSPACE: the workspace bytes the failed call needs
 * (thread-local).  am_sdf_bwd's need depends on the data (its number of
 * table-gradient pairs), so a caller may pass less than the worst-case size
 * query, keep a grow-only workspace and call again with this many bytes. */
size_t am_last_workspace_need(void);
int am_version(void);

/* Kernel timers for measurement: when enabled, the named hot kernels are
 * bracketed by HIP events on their launch stream.  am_timer_read
 * synchronises on the recorded events and returns the summed duration and
 * launch count of one kernel name (e.g. "geo_bwd").  Not thread-safe;
 * meant for benchmarks.  Enabling resets all counters. */
int am_timer_enable(int32_t on);
int am_timer_read(const char* name, double* total_ms, int64_t* count);
/* Work units recorded under `name` since am_timer_enable (e.g.
 * "table_pairs": the (row, value) pairs the table-gradient sorts received). */
int am_timer_units(const char* name, int64_t* units);

/* Implementation variant switches, for measurement and tests (no
 * counterpart in the reference).  "table_prepartition" (default 0): 1 makes
 * the SDF backward's pair pass write its (row, value) pairs partitioned by
 * bits 9-16 of their row, so the table-gradient sort makes one pass over the
 * remaining bits instead of two (it measured slower: scattered stores); 0:
 * compact per-tile pairs and two sort passes.  Both give bit-identical gradients.
 * "sdf_bwd_sync_free" (default 0): 1 makes am_sdf_bwd size its pair arrays
 * for the worst-case pair count instead of reading the exact count back to
 * the host (no host wait; the sort covers the bound, so it is slower; the
 * workspace must then hold am_sdf_bwd_workspace_size(), else AM_ERR_WORKSPACE
 * with that need); 0: one host read of the pair count.  "table_atomic"
 * (default 0): 1 makes am_sdf_bwd add its merged pairs into the table
 * gradient with float atomics (no sort, no host read; the summation order
 * varies run to run; measured 3x slower than the sort on a 0.87 M-sample
 * step).  Unknown names return
 * AM_ERR_INVALID. */
int am_set_flag(const char* name, int32_t value);

/* ------------------------------------------------------------------------ */
/* Multiresolution hash grid (tcnn HashGrid through ProgressiveBandHashGrid,  */
/* models/network_utils.py:40-66, and get_encoding, :83-93).                 */
/* ------------------------------------------------------------------------ */
typedef struct am_grid_desc {
    int32_t n_levels;           /* L <= 32 */
    int32_t n_features;         /* F in {1, 2, 4, 8} */
    int32_t log2_hashmap_size;  /* T = 2^log2 entries per hashed level */
    int32_t base_resolution;
    float per_level_scale;
    uint32_t n_entries;         /* total table rows (sum of level sizes) */
    uint32_t hashed_mask;       /* bit l set: level l uses the prime hash */
    float scale[AM_MAX_LEVELS]; /* fmaf(scale, x, 0.5) per level */
    uint32_t res[AM_MAX_LEVELS];
    uint32_t offset[AM_MAX_LEVELS]; /* first table row of level l */
    uint32_t size[AM_MAX_LEVELS];   /* rows of level l (the index modulus) */
} am_grid_desc;

/* Fills the level table by tcnn's published rule (scale_l = 2^(l*log2(s))*base
 * - 1 in fp32, res_l = ceil(scale_l)+1, rows = min(next_mult8(res^3), T)).
 * res_override (nullable, L entries) replaces the computed resolutions (the
 * scale is then res-1), for matching another implementation's table.
 * Replaces tcnn.Encoding construction at models/network_utils.py:45-47. */
int am_grid_init(am_grid_desc* g, int32_t n_levels, int32_t n_features,
                 int32_t log2_hashmap_size, int32_t base_resolution,
                 float per_level_scale, const uint32_t* res_override);

/* Encode x[n,3] in [0,1] -> out[n, L*F] level-major (column l*F+f); levels
 * >= active_levels give 0 (the ProgressiveBandHashGrid mask,
 * models/network_utils.py:53-58).  table: [n_entries, F] fp32.
 * Replaces ProgressiveBandHashGrid.forward (models/network_utils.py:56-59). */
int am_hashgrid_fwd(const am_grid_desc* g, const float* x, int64_t n, const float* table,
                    int32_t active_levels, float* out, void* stream);

/* Deterministic backward: dtable[n_entries,F] (fully overwritten) by
 * sort-by-row + segmented reduce; dx[n,3] (nullable) = dL/dx.
 * Replaces tcnn's encoding backward (atomic scatter) behind the same call. */
size_t am_hashgrid_bwd_workspace_size(const am_grid_desc* g, int64_t n, int32_t active_levels);
int am_hashgrid_bwd(const am_grid_desc* g, const float* x, int64_t n, const float* table,
                    const float* dy, int32_t active_levels, float* dtable, float* dx,
                    void* workspace, size_t workspace_bytes, void* stream);
/* am_hashgrid_fwd / am_hashgrid_bwd evaluating the points in the order
 * `order` (a permutation of [0, n); point k of the launch is the caller's
 * point order[k]); x, out, dy and dx stay in the caller's order.  No
 * counterpart in the reference: the binding passes the Morton order of large
 * batches so that nearby points gather nearby rows.  Outputs are identical;
 * the table gradient's fp32 sums run in another fixed order. */
int am_hashgrid_fwd_ordered(const am_grid_desc* g, const int32_t* order, const float* x,
                            int64_t n, const float* table, int32_t active_levels, float* out,
                            void* stream);
int am_hashgrid_bwd_ordered(const am_grid_desc* g, const int32_t* order, const float* x,
                            int64_t n, const float* table, const float* dy,
                            int32_t active_levels, float* dtable, float* dx, void* workspace,
                            size_t workspace_bytes, void* stream);

/* ------------------------------------------------------------------------ */
/* VanillaMLP (models/network_utils.py:96-140, get_mlp :177-185).            */
/* Parameters: one flat fp32 buffer [W0,b0,W1,b1,...,Wlast,blast], each W    */
/* row-major [out,in] as nn.Linear.weight (weight-norm already applied by    */
/* the host: W = g*v/||v|| per output row, :132-133).                        */
/* ------------------------------------------------------------------------ */
#define AM_ACT_NONE 0
#define AM_ACT_RELU 1
#define AM_ACT_SOFTPLUS 2 /* Softplus(beta), threshold 20 (torch default) */
#define AM_ACT_SIGMOID 3

#define AM_PREC_FP32 0
#define AM_PREC_MIXED 1

/* VolumeSDF grad_type (models/geometry.py:214-234) */
#define AM_GRAD_FD 0       /* 'finite_difference': 6 central differences */
#define AM_GRAD_ANALYTIC 1 /* 'analytic': d sdf / d x through the encoding and
                              the MLP (one or two hidden layers), backward
                              with the second-order terms of the gradient
                              output */

typedef struct am_mlp_desc {
    int32_t n_in;            /* <= 64 (stand-alone MLP); the fused SDF MLP takes
                                3 + 2 L, the fused radiance MLP <= 96 */
    int32_t n_out;           /* <= 64 (stand-alone MLP); the fused SDF MLP takes
                                <= 16, or <= 80 with two hidden layers (the 65
                                outputs of the Neuralangelo configurations) */
    int32_t n_hidden_layers; /* 1..4 */
    int32_t n_neurons;       /* 64 */
    int32_t hidden_act;      /* AM_ACT_RELU or AM_ACT_SOFTPLUS */
    int32_t output_act;      /* AM_ACT_NONE / RELU / SIGMOID */
    float softplus_beta;
    int32_t precision;       /* AM_PREC_FP32, or AM_PREC_MIXED: matmuls on bf16
                                MFMA with fp32 accumulation (the fused radiance
                                kernels; the SDF MLP and the stand-alone MLP
                                take fp32 only) */
} am_mlp_desc;

int64_t am_mlp_n_params(const am_mlp_desc* m);
int am_mlp_fwd(const am_mlp_desc* m, const float* x, int64_t n, const float* params, float* y,
               void* stream);
size_t am_mlp_bwd_workspace_size(const am_mlp_desc* m, int64_t n);
/* dx (nullable) [n,n_in]; dparams [n_params] fully overwritten (deterministic). */
int am_mlp_bwd(const am_mlp_desc* m, const float* x, int64_t n, const float* params,
               const float* dy, float* dx, float* dparams, void* workspace,
               size_t workspace_bytes, void* stream);

/* Real spherical harmonics of a direction given in [0,1]^3 (tcnn
 * SphericalHarmonics, models/texture.py:17,24-25): out[n, degree^2],
 * degree 1..4.  Replaces the tcnn encoding call at models/texture.py:25. */
int am_sh_fwd(const float* dirs01, int64_t n, int32_t degree, float* out, void* stream);

/* ------------------------------------------------------------------------ */
/* Ray marching (nerfacc 0.3.3 ray_aabb_intersect + ray_marching as called at */
/* models/neus.py:209-220).  rays: [R,6] = origin, unit direction.           */
/* ------------------------------------------------------------------------ */
int am_ray_aabb(const float* rays, int64_t n_rays, const float* aabb6_host, float* t_near,
                float* t_far, void* stream);

/* Pass 1: packed[R,2] = (start, count) and total[1] (device int32).
 * jitter (nullable [R]): stratified start t_near + u*step.
 * grid (nullable, uint8 [res^3], x-major): occupancy; NULL = all occupied.
 * cap: per-ray sample cap (0 = none).  roi6_host: grid AABB. */
size_t am_march_workspace_size(int64_t n_rays);
int am_march_count(const float* rays, int64_t n_rays, const float* t_near, const float* t_far,
                   const float* jitter, const float* roi6_host, const uint8_t* grid,
                   int32_t grid_res, float step, int32_t cap, int32_t* packed,
                   int32_t* total, void* workspace, size_t workspace_bytes, void* stream);
/* Pass 2: ray-sorted samples; ray_indices[N] int32, t_starts/t_ends[N]. */
int am_march_fill(const float* rays, int64_t n_rays, const float* t_near, const float* t_far,
                  const float* jitter, const float* roi6_host, const uint8_t* grid,
                  int32_t grid_res, float step, int32_t cap, const int32_t* packed,
                  int32_t* ray_indices, float* t_starts, float* t_ends, void* stream);

/* Staged form of the two passes, for a per-ray sample cap > 0: the count
 * pass also writes each ray's samples to a cap-strided area of the workspace
 * (am_march_staged_workspace_size(n_rays, cap) bytes), and the fill pass
 * copies them to the packed outputs instead of marching again.  Same
 * arguments and outputs as am_march_count / am_march_fill, bit-identical
 * samples; the fill call takes the workspace the count call filled. */
size_t am_march_staged_workspace_size(int64_t n_rays, int32_t cap);
int am_march_count_staged(const float* rays, int64_t n_rays, const float* t_near,
                          const float* t_far, const float* jitter, const float* roi6_host,
                          const uint8_t* grid, int32_t grid_res, float step, int32_t cap,
                          int32_t* packed, int32_t* total, void* workspace,
                          size_t workspace_bytes, void* stream);
int am_march_fill_staged(int64_t n_rays, int32_t cap, const int32_t* packed,
                         const void* workspace, size_t workspace_bytes, int32_t* ray_indices,
                         float* t_starts, float* t_ends, void* stream);

/* Background marching: nerfacc ray_marching with ContractionType.
 * UN_BOUNDED_SPHERE and cone angle as called by forward_bg_ (models/neus.py:
 * 141-171, scene_aabb None): per ray [t_near, t_far] are given (near plane,
 * far plane and stratified jitter applied by the caller); steps of
 * dt = clamp(t0 * cone_angle, step, 1e10); a sample is kept when its
 * midpoint's cell of the contracted grid (res^3 over roi6, or NULL: all) is
 * occupied.  Same two passes and outputs as am_march_count / am_march_fill. */
int am_march_bg_count(const float* rays, int64_t n_rays, const float* t_near, const float* t_far,
                      const float* roi6_host, const uint8_t* grid, int32_t grid_res, float step,
                      float cone_angle, int32_t cap, int32_t* packed, int32_t* total,
                      void* workspace, size_t workspace_bytes, void* stream);
int am_march_bg_fill(const float* rays, int64_t n_rays, const float* t_near, const float* t_far,
                     const float* roi6_host, const uint8_t* grid, int32_t grid_res, float step,
                     float cone_angle, int32_t cap, const int32_t* packed, int32_t* ray_indices,
                     float* t_starts, float* t_ends, void* stream);
/* Early-stop filter of the background samples (nerfacc 0.3.3 ray_marching
 * with sigma_fn and alpha_thre 0, models/neus.py:141-167): with the samples'
 * density sigma[n] (packed per ray), keeps each ray's samples while the
 * transmittance prod_{i<j} (1 - alpha_i), alpha = 1 - exp(-sigma dt), is at
 * least early_stop_eps (1e-4 in the reference).  The kept samples are a
 * prefix of each ray's: packed_out[R,2] = (start, kept count), total[1];
 * am_march_bg_fill with packed_out then writes exactly them.  Workspace:
 * am_march_workspace_size(n_rays). */
int am_march_bg_visible(const float* sigma, const float* t_starts, const float* t_ends,
                        const int32_t* packed, int64_t n_rays, float early_stop_eps,
                        int32_t* packed_out, int32_t* total, void* workspace,
                        size_t workspace_bytes, void* stream);

/* Background compositing (nerfacc render_weight_from_density +
 * accumulate_along_rays, models/neus.py:174-178): weights[n] =
 * exp(-sum_{j<i} s_j) (1 - exp(-s_i)), s = density (t_end - t_start);
 * trans[n] (kept for the backward); opacity[R], depth[R] (midpoints),
 * comp_rgb[R,3].  Backward to density[n] and rgb[n,3] (upstream gradients
 * nullable = 0). */
int am_composite_density_fwd(const float* density, const float* t_starts, const float* t_ends,
                             const float* rgb, const int32_t* packed, int64_t n_rays, int64_t n,
                             float* weights, float* trans, float* opacity, float* depth,
                             float* comp_rgb, void* stream);
int am_composite_density_bwd(const float* density, const float* t_starts, const float* t_ends,
                             const float* rgb, const int32_t* packed, int64_t n_rays, int64_t n,
                             const float* weights, const float* trans, const float* g_weights,
                             const float* g_opacity, const float* g_depth,
                             const float* g_comp_rgb, float* g_density, float* g_rgb,
                             void* stream);

/* Training-ray sampling (systems/neus.py:35-118, train stage; get_rays at
 * models/ray_utils.py:23-31): for ray k, view v = index[k], pixel (ys[k],
 * xs[k]): rays[k] = (c2w[v][:, 3], normalize(c2w[v][:3, :3] d)) with d =
 * directions[y, x] (or directions[v, y, x] when dirs_per_view), the rotation
 * summed left to right and F.normalize's max(|r|, 1e-12); rgb[k] =
 * images[v, y, x, 0:3] (C channels per pixel), fg_mask[k] = fg_masks[v, y, x]
 * (1 when NULL); apply_mask blends rgb * fg + bg_color * (1 - fg) with the
 * device colour bg_color[3].  c2w is [n_views, 3, 4].  Draws out of range are
 * clamped and counted into *n_out_of_range (device int32, not cleared). */
int am_sample_rays(const float* images, const float* fg_masks, const float* c2w,
                   const float* directions, int32_t n_views, int32_t H, int32_t W, int32_t C,
                   int32_t dirs_per_view, const float* bg_color, int32_t apply_mask,
                   const int32_t* index, const int32_t* ys, const int32_t* xs, int64_t n,
                   float* rays, float* rgb, float* fg_mask, int32_t* n_out_of_range,
                   void* stream);

/* Spatial (Morton) order key of each point of [-radius, radius]^3: 10 bits
 * per axis, interleaved, x highest (points outside are clamped to the box).
 * No counterpart in the reference: the fused SDF field processes its samples
 * in this order (a stable sort of the keys) so that concurrently running
 * waves gather nearby table rows; outputs are returned in the caller's order. */
int am_morton_keys(const float* pts, int64_t n, float radius, int32_t* keys, void* stream);
/* The order in one call, with `bits` (1..10) bits per axis: perm[n] = the
 * points sorted by key (stable: points of one cell keep the caller's order),
 * inv[n] its inverse (inv[perm[i]] = i).  Workspace:
 * am_morton_order_workspace_size(n, bits). */
size_t am_morton_order_workspace_size(int64_t n, int32_t bits);
int am_morton_order(const float* pts, int64_t n, float radius, int32_t bits, int32_t* perm,
                    int32_t* inv, void* workspace, size_t workspace_bytes, void* stream);

/* Sample positions for models/neus.py:222-227: pts = o[idx] + d[idx]*mid,
 * mid = (t0+t1)/2, dists = t1-t0 (each a separate fp32 op, as torch). */
int am_sample_points(const float* rays, const int32_t* ray_indices, const float* t_starts,
                     const float* t_ends, int64_t n, float* pts, float* mid, float* dists,
                     void* stream);

/* Unbounded-sphere contraction of VolumeDensity: contract_to_unisphere
 * (models/geometry.py:19-31) with ContractionType.UN_BOUNDED_SPHERE, as
 * VolumeDensity.forward calls it (:160): x [n,3] world points -> out [n,3]
 * in [0,1]^3 (cube [-r,r]^3 -> [-1,1]^3; norm m > 1 -> (2 - 1/m) x/m; then
 * x/4 + 1/2).  _bwd: g_x = J^T g_out at the same points. */
int am_contract_unbounded(const float* x, int64_t n, float radius, float* out, void* stream);
int am_contract_unbounded_bwd(const float* x, const float* g_out, int64_t n, float radius,
                              float* g_x, void* stream);

/* ------------------------------------------------------------------------ */
/* Fused VolumeSDF.forward (models/geometry.py:195-287): contraction,         */
/* encode + SDF MLP of the base point, 6 central-difference points (grad)    */
/* and, with laplace, the 6+ curvature points around the tangent-shifted     */
/* point, all in one launch.  params: flat SDF-MLP buffer (am_mlp_desc with   */
/* n_in = 3 + L*F, n_out = feature_dim).                                      */
/* ------------------------------------------------------------------------ */
typedef struct am_sdf_desc {
    float radius;         /* scene radius r: AABB [-r,r]^3 */
    float fd_eps;         /* finite-difference eps (VolumeSDF._finite_difference_eps) */
    int32_t with_grad;    /* 6 central-difference evaluations */
    int32_t with_laplace; /* 6 more around the shifted point (needs with_grad) */
    int32_t recontract;   /* 1: shifted point re-contracted as the reference does
                             (models/geometry.py:246,264-265); 0: shifted in world space */
    int32_t active_levels;
    int32_t precision;    /* AM_PREC_FP32, or AM_PREC_MIXED: table rows stored
                             in fp16 ([n_entries,2] __half) and table-gradient
                             pair values in fp16 (expects loss-scaled upstream
                             gradients); blends, sums and dtable stay fp32.
                             The SDF MLP descriptor must be AM_PREC_FP32 in
                             both forms: the central differences of the FD
                             normal divide each evaluation's rounding by eps */
    int32_t grad_type;    /* AM_GRAD_FD or AM_GRAD_ANALYTIC (with_grad) */
} am_sdf_desc;

/* Outputs (each nullable except sdf): sdf[n], grad[n,3], normal[n,3] =
 * F.normalize(grad), feature[n, n_out+3] = cat(mlp_out, x01*2-1),
 * laplace[n], fd_values[n,12] (the 6 central-difference SDF values, or the
 * analytic grad in [0,3), then the 6 curvature SDF values, kept for the
 * backward), enc_store[am_sdf_store_floats(n)] (the
 * encodings of every evaluation and the position derivatives of the
 * curvature evaluations, kept so the backward does not gather the table
 * again).  randn[n,3]: the curvature directions (torch.randn_like). */
int am_sdf_fwd(const am_grid_desc* g, const am_mlp_desc* m, const am_sdf_desc* s,
               const float* pts, const float* randn, int64_t n, const void* table,
               const float* params, float* sdf, float* grad, float* normal, float* feature,
               float* laplace, float* fd_values, float* enc_store, void* stream);
/* Floats of enc_store for n samples (-1 on invalid descriptors). */
int64_t am_sdf_store_floats(const am_grid_desc* g, const am_mlp_desc* m, const am_sdf_desc* s,
                            int64_t n);
/* Worst-case bound: every sample's every corner a pair of its own. */
size_t am_sdf_bwd_workspace_size(const am_grid_desc* g, const am_mlp_desc* m,
                                 const am_sdf_desc* s, int64_t n);
/* Upstream gradients are nullable (= 0).  fd_values: from the forward
 * (nullable: recomputed).  enc_store: from the forward (nullable: the
 * backward re-gathers the table, the memory-lean "recompute" form; needs
 * fd_values when given).  dtable [n_entries,F] and dparams are fully
 * overwritten; both reductions are deterministic.  The call counts the
 * table-gradient pairs first and reads the count back to the host (a stream
 * sync), carving the pair arrays for exactly that many; a workspace below
 * that need returns AM_ERR_WORKSPACE before any output is written
 * (am_last_workspace_need).  The emit pass's consistency flag (its pairs
 * equal the count pass's) is copied to pinned host memory without a wait and
 * read by later calls, after their own count sync, and by am_sdf_bwd_check;
 * a set flag returns AM_ERR_INTERNAL there.
 * A workspace of at least am_sdf_bwd_workspace_size makes the call free of
 * host reads: the pair arrays take the worst-case bound, their unused tail
 * sentinel keys, and the sort covers the bound (more sort work, identical
 * gradients).
 * Replaces the autograd backward of VolumeSDF.forward
 * (models/geometry.py:195-287) through tcnn's encoding backward. */
/* Waits for the consistency flags of every am_sdf_bwd call so far:
 * AM_OK, or AM_ERR_INTERNAL if one was set.  No counterpart in the
 * reference (its backward has no pair passes to cross-check). */
int am_sdf_bwd_check(void);
int am_sdf_bwd(const am_grid_desc* g, const am_mlp_desc* m, const am_sdf_desc* s,
               const float* pts, const float* randn, int64_t n, const void* table,
               const float* params, const float* fd_values, const float* enc_store,
               const float* g_sdf, const float* g_grad, const float* g_normal,
               const float* g_feature, const float* g_laplace, float* dtable, float* dparams,
               void* workspace, size_t workspace_bytes, void* stream);
/* The same two calls evaluating the samples in the order `order` (a
 * permutation of [0, n): kernel sample i is the caller's sample order[i],
 * for instance from am_morton_order).  Positions, directions, outputs and
 * upstream gradients stay in the caller's order; fd_values and enc_store
 * are in the evaluation order, so the backward takes the forward's order.
 * No counterpart in the reference (it evaluates in ray order); per-sample
 * results are identical, only the gradients' fp32 summation order differs. */
int am_sdf_fwd_ordered(const am_grid_desc* g, const am_mlp_desc* m, const am_sdf_desc* s,
                       const int32_t* order, const float* pts, const float* randn, int64_t n,
                       const void* table, const float* params, float* sdf, float* grad,
                       float* normal, float* feature, float* laplace, float* fd_values,
                       float* enc_store, void* stream);
int am_sdf_bwd_ordered(const am_grid_desc* g, const am_mlp_desc* m, const am_sdf_desc* s,
                       const int32_t* order, const float* pts, const float* randn, int64_t n,
                       const void* table, const float* params, const float* fd_values,
                       const float* enc_store, const float* g_sdf, const float* g_grad,
                       const float* g_normal, const float* g_feature, const float* g_laplace,
                       float* dtable, float* dparams, void* workspace, size_t workspace_bytes,
                       void* stream);

/* ------------------------------------------------------------------------ */
/* Fused VolumeRadiance.forward (models/texture.py:23-30): SH(deg) of        */
/* (d+1)/2, cat(feature, sh, normal) -> MLP -> sigmoid.  dirs are gathered    */
/* from rays[ray_indices] (models/neus.py:224).  base_color = 1:              */
/* VolumeDualColor.forward (models/texture.py:38-55) adds sigmoid(feature[:,  */
/* 1:4]) (feature_dim >= 4).                                                  */
/* ------------------------------------------------------------------------ */
int am_radiance_fwd(const am_mlp_desc* m, int32_t sh_degree, int32_t base_color,
                    const float* feature, int32_t feature_dim, const float* rays,
                    const int32_t* ray_indices, const float* normal, int64_t n,
                    const float* params, float* rgb, void* stream);
size_t am_radiance_bwd_workspace_size(const am_mlp_desc* m, int64_t n);
int am_radiance_bwd(const am_mlp_desc* m, int32_t sh_degree, int32_t base_color,
                    const float* feature, int32_t feature_dim, const float* rays,
                    const int32_t* ray_indices,
                    const float* normal, int64_t n, const float* params, const float* g_rgb,
                    float* g_feature, float* g_normal, float* dparams, void* workspace,
                    size_t workspace_bytes, void* stream);

/* ------------------------------------------------------------------------ */
/* NeuS alpha + transmittance + accumulation (models/neus.py:117-139,231-241 */
/* with nerfacc render_weight_from_alpha / accumulate_along_rays).           */
/* packed[R,2] (start,count) over ray-sorted samples.  inv_s: device scalar. */
/* ------------------------------------------------------------------------ */
typedef struct am_composite_desc {
    float cos_anneal_ratio;
    int32_t kahan; /* 1: compensated transmittance/accumulation sums */
} am_composite_desc;

/* Outputs: weights[n] (w_i = T_i*alpha_i), trans[n] (T_i, nullable, kept for
 * backward), opacity[R], depth[R], comp_rgb[R,3], comp_normal[R,3]
 * (normalised, 0 where opacity <= 0.1), comp_normal_raw[R,3] (the plain sum,
 * nullable, kept for backward).  Rays without samples give zeros. */
int am_composite_fwd(const am_composite_desc* c, const float* sdf, const float* normal,
                     const float* rays, const int32_t* ray_indices, const float* t_starts,
                     const float* t_ends, const float* rgb, const float* inv_s,
                     const int32_t* packed, int64_t n_rays, int64_t n, float* weights,
                     float* trans, float* opacity, float* depth, float* comp_rgb,
                     float* comp_normal, float* comp_normal_raw, void* stream);
size_t am_composite_bwd_workspace_size(int64_t n_rays, int64_t n);
/* Upstream grads nullable (= 0).  trans, opacity, comp_normal_raw: from the
 * forward.  g_inv_s: device scalar, a deterministic reduction. */
int am_composite_bwd(const am_composite_desc* c, const float* sdf, const float* normal,
                     const float* rays, const int32_t* ray_indices, const float* t_starts,
                     const float* t_ends, const float* rgb, const float* inv_s,
                     const int32_t* packed, int64_t n_rays, int64_t n, const float* trans,
                     const float* opacity, const float* comp_normal_raw,
                     const float* g_weights, const float* g_opacity, const float* g_depth,
                     const float* g_comp_rgb, const float* g_comp_normal, float* g_sdf,
                     float* g_normal, float* g_rgb, float* g_inv_s, void* workspace,
                     size_t workspace_bytes, void* stream);

/* am_composite_bwd without a workspace argument, for callers that cannot
 * allocate device memory (code under HIP graph capture, serving loops,
 * harnesses that forbid hipMalloc).  No counterpart in the reference.  Its
 * scratch is a static device array of the library (about 4.1 MiB per device,
 * held until the process ends): n_rays <= 1,048,576, larger calls return
 * AM_ERR_INVALID before any launch.  Same arguments (less the workspace) and
 * bit-identical outputs.  The array is shared by every call on a device:
 * calls that overlap on one device race; the stream must belong to the
 * current device. */
int am_composite_bwd_noalloc(const am_composite_desc* c, const float* sdf, const float* normal,
                             const float* rays, const int32_t* ray_indices,
                             const float* t_starts, const float* t_ends, const float* rgb,
                             const float* inv_s, const int32_t* packed, int64_t n_rays,
                             int64_t n, const float* trans, const float* opacity,
                             const float* comp_normal_raw, const float* g_weights,
                             const float* g_opacity, const float* g_depth,
                             const float* g_comp_rgb, const float* g_comp_normal, float* g_sdf,
                             float* g_normal, float* g_rgb, float* g_inv_s, void* stream);

/* ------------------------------------------------------------------------ */
/* Occupancy grid (nerfacc 0.3.3 OccupancyGrid, models/neus.py:63-68,90-111). */
/* ------------------------------------------------------------------------ */
/* EMA update of occs[cells] with the NeuS alpha estimate of the SDF at the   */
/* jittered cell points (occ_eval_fn, models/neus.py:90-101), then the      */
/* binary threshold min(mean(occs), thre).  cells (nullable = all cells in  */
/* order) [m] int32; jitter [m,3] in [0,1).                                   */
size_t am_occ_update_workspace_size(int32_t grid_res, int64_t m);
int am_occ_update(const am_grid_desc* g, const am_mlp_desc* mlp, const am_sdf_desc* s,
                  const void* table, const float* params, const float* inv_s,
                  float step, const int32_t* cells, const float* jitter, int64_t m,
                  int32_t grid_res, float ema_decay, float occ_thre, float* occs,
                  uint8_t* binary, void* workspace, size_t workspace_bytes, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* ANGELO_AMD_H */
