/*
 * strns2cards_hip.h — C ABI of the MI355X (gfx950) HIP library behind
 * strns2cards_amd.  Plain pointers and sizes only: every array argument is a
 * device pointer (HIP device memory, C-contiguous), every size an int, and
 * `stream` is a hipStream_t passed as void* (NULL = the legacy default stream).
 *
 * Each entry point below replaces one launcher of the reference's
 * `strns2cards::cuda` namespace, declared at
 * pystrns2cards/src/pystrns2cards.cpp:13-192 (kenji-tojo/strns2cards) and
 * bound to Python by the nanobind `_core` module at
 * pystrns2cards/src/pystrns2cards.cpp:204-924.  Argument order and meaning
 * are the reference's; two things differ on purpose:
 *   - every function returns a status (S2C_OK, a negative S2C_ERR_*, or a
 *     positive hipError_t) instead of calling exit() on a CUDA error
 *     (cuda_common.cuh:11-14); s2c_last_error() gives the message;
 *   - launchers that return a count in the reference write it through a
 *     trailing `int* out` host pointer (the call then synchronizes `stream`,
 *     as the reference's cudaMemcpy does).
 * Temporaries are taken from HIP's stream-ordered allocator on `stream`.
 * List outputs are produced in the stable (primitive-major) order.
 */
#ifndef STRNS2CARDS_HIP_H
#define STRNS2CARDS_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

#define S2C_OK 0
#define S2C_ERR_ARG (-1)      /* invalid argument (shape, size, null) */
#define S2C_ERR_OVERFLOW (-2) /* a count exceeds INT32_MAX */

/* Library version as major*10000 + minor*100 + patch. */
int s2c_version(void);
/* Message of the last failing call on this host thread ("" if none). */
const char* s2c_last_error(void);
/* Hash of the sources the library was built from (build.py:source_hash):
 * the package refuses a library whose hash is not that of the sources
 * beside it.  No counterpart in the reference. */
const char* s2c_build_id(void);

/* ---- triangle rasterization: rasterize.cu -------------------------------- */

/* rasterize.cu:61-88 (compute_triangle_rects). Writes rects (B*T,4) =
 * [h0,h_len,w0,w_len] and the inclusive prefix of h_len*w_len; *num_frags_out
 * receives the candidate total (0 when B*T == 0). */
int s2c_compute_triangle_rects(int H, int W, int B, int V, const float* pos,
                               int T, const int* tri, int* triangle_rects,
                               int* frag_prefix_sum, int* num_frags_out,
                               void* stream);

/* rasterize.cu:174-195 (compute_fragments). Candidates that fail the
 * coverage / depth tests are not written (the caller pre-fills frag_pix with
 * -1, rasterize.py:179). */
int s2c_compute_fragments(int H, int W, int V, const float* pos, int T,
                          const int* tri, int num_tris, int num_frags,
                          const int* frag_prefix_sum, const int* triangle_rects,
                          int* frag_pix, float* frag_attrs, void* stream);

/* rasterize.cu:258-278 (depth_test). rast_out (B,H,W,4) is zero-filled by
 * the caller; key = (bits(z+2) << 32) | fragment index, smallest wins. */
int s2c_depth_test(int B, int H, int W, int num_fragments, const int* frag_pix,
                   const float* frag_attrs, float* rast_out, void* stream);

/* s2c_depth_test without any allocation.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under HIP graph
 * capture, serving loops, harnesses that forbid hipMalloc).  Same arguments,
 * result and errors as s2c_depth_test; its key buffer is a static device
 * array of 4,194,304 int64 (32 MiB) per GPU, held until the process ends, so
 * B*H*W above 4,194,304 returns S2C_ERR_ARG.  Calls on one GPU share that
 * buffer: calls that overlap on one GPU (other streams or threads) race. */
int s2c_depth_test_noalloc(int B, int H, int W, int num_fragments,
                           const int* frag_pix, const float* frag_attrs,
                           float* rast_out, void* stream);

/* rasterize.cu:336-365 (filter_valid_fragments), stable order. */
int s2c_filter_valid_fragments(int num_frags, const int* frag_pix,
                               const float* frag_attrs, int* frag_pix_out,
                               float* frag_attrs_out, int* valid_count_out,
                               void* stream);

/* rasterize.cu:440-455 (interpolate_triangle_attributes). */
int s2c_interpolate_triangle_attributes(int B, int H, int W, int attr_dim,
                                        const float* rast, const float* attr,
                                        const int* tri, float* image,
                                        void* stream);

/* rasterize.cu:457-472 (backward_interpolate_triangle_attributes);
 * accumulates into d_attr. */
int s2c_backward_interpolate_triangle_attributes(int B, int H, int W,
                                                 int attr_dim,
                                                 const float* rast,
                                                 const float* d_image,
                                                 const int* tri, float* d_attr,
                                                 void* stream);

/* rasterize.cu:500-520 (cluster_mask_from_fragments); ORs into bitset. */
int s2c_cluster_mask_from_fragments(int num_frags, const int* frag_pix,
                                    const float* frag_attrs,
                                    const int* prim2cluster, int H, int W,
                                    int num_slots, int* bitset, void* stream);

/* rasterize.py:170-214 (a zeroed (B, H, W, num_slots) bitset, then
 * cluster_mask_from_fragments) as one binned pass that writes every word
 * of `bitset` (no clear, no global OR per fragment).  Fragments with a prim
 * id <= 0, a pixel outside B x H x W or a cluster >= 32 * num_slots set no
 * bit.  No host wait; the binning workspace is that of
 * s2c_dda_cluster_mask on the same stream. */
int s2c_fragment_cluster_mask(int B, int H, int W, int num_frags,
                              const int* frag_pix, const float* frag_attrs,
                              const int* prim2cluster, int num_slots,
                              int* bitset, void* stream);

/* max over i of frag_pix[i][0] (rasterize.py:197: the view count B is that
 * plus one) into *max_out; num_frags > 0, frag_pix 16-byte aligned.
 * Synchronizes the stream. */
int s2c_max_frag_view(int num_frags, const int* frag_pix, int* max_out,
                      void* stream);

/* ---- bitset utilities: bitset_utils.cu ----------------------------------- */

/* bitset_utils.cu:55-68 (accumulate_bitset_rgb). */
int s2c_accumulate_bitset_rgb(int B, int H, int W, int num_slots,
                              const int* bitset, const float* cluster_rgb,
                              float* accum_color, void* stream);

/* bitset_utils.cu:70-84 (popcount_bitset). */
int s2c_popcount_bitset(int B, int H, int W, int num_slots, const int* bitset,
                        float* count_image, void* stream);

/* ---- DDA line rasterization: dda.cu --------------------------------------- */

/* dda.cu:89-122 (dda_compute_span). edge_ids[i] = b*E + e. */
int s2c_dda_compute_span(int num_prims, const int* edge_ids, int H, int W,
                         int V, const float* pos, int E, const int* edges,
                         int* frag_prefix_sum, float* frag_slopes,
                         float* frag_spans, int* num_frags_out, void* stream);

/* dda.cu:227-256 (dda_compute_fragments). Failing candidates not written. */
int s2c_dda_compute_fragments(int num_prims, int num_frags,
                              const int* frag_prefix_sum, const int* edge_ids,
                              const float* frag_slopes, const float* frag_spans,
                              int H, int W, int V, const float* pos, int E,
                              const int* edges, int* frag_pix,
                              float* frag_attrs, void* stream);

/* dda.py:5-76 (dda_compute_fragments: span, candidate expansion and
 * filter_valid_fragments) as two calls in the space of surviving fragments,
 * with no candidate list.  s2c_dda_count_fragments writes the inclusive
 * prefix of each primitive's surviving fragments to valid_prefix
 * (num_prims ints) and their total; s2c_dda_emit_fragments then writes the
 * num_frags survivors, in the reference's compacted order, to frag_pix
 * (num_frags, 3) and frag_attrs (num_frags, 4).  edge_ids may be null
 * (primitive p is edge p, the reference's default arange). */
int s2c_dda_count_fragments(int num_prims, const int* edge_ids, int H, int W,
                            int V, const float* pos, int E, const int* edges,
                            int* valid_prefix, int* num_frags_out,
                            void* stream);
int s2c_dda_emit_fragments(int num_prims, int num_frags,
                           const int* valid_prefix, const int* edge_ids, int H,
                           int W, int V, const float* pos, int E,
                           const int* edges, int* frag_pix, float* frag_attrs,
                           void* stream);

/* dda.cu:286-298 (dda_interpolate_attributes). */
int s2c_dda_interpolate_attributes(int B, int H, int W, int C,
                                   const float* rast_dda, const float* attr,
                                   const int* edges, float* image,
                                   void* stream);

/* dda.cu:329-342 (backward_dda_interpolate_attributes); accumulates. */
int s2c_backward_dda_interpolate_attributes(int B, int H, int W,
                                            const float* rast_dda, int C,
                                            const float* d_image,
                                            const int* edges, float* d_attr,
                                            void* stream);

/* ---- antialias: antialias.cu, compact_utils.cu ---------------------------- */

/* antialias.cu:97-119 (mark_discontinuity_edges). prim_ids (B*E) is
 * -1-filled and normals (B*E,2) zero-filled by the caller. */
int s2c_mark_discontinuity_edges(int H, int W, int B, int V, const float* pos,
                                 const int* tri, int E, const int* edges,
                                 const int* edge2tri, int* prim_ids,
                                 float* normals, void* stream);

/* antialias.cu:282-317 (backward_antialiased_cluster_bitset); accumulates
 * into d_pos (B,V,4). */
int s2c_backward_antialiased_cluster_bitset(
    int num_frags, const int* frag_pix, const float* frag_attrs_dda, int H,
    int W, int num_slots, const int* bitset, const int* target_bitset, int V,
    const float* pos, int E, const int* edges, const float* normals,
    const int* edge2cluster, float* d_pos, float kernel_radius, float rho,
    void* stream);

/* s2c_backward_antialiased_cluster_bitset given the bitsets' view count B
 * (num_views; 0 = unknown).  No reference launcher corresponds: the
 * reference's takes no B.  With it, and num_slots >= 16, the residual bits
 * are read from word planes (B, S, H, W) of the two bitsets built by a
 * coalesced transpose into a temporary of B*H*W*S*8 bytes, instead of one
 * word from each pixel's S-word record; fragments of views >= B add
 * nothing.  Same result, up to the order of the d_pos atomics. */
int s2c_backward_antialiased_cluster_bitset_views(
    int num_views, int num_frags, const int* frag_pix, const float* frag_attrs_dda,
    int H, int W, int num_slots, const int* bitset, const int* target_bitset, int V,
    const float* pos, int E, const int* edges, const float* normals,
    const int* edge2cluster, float* d_pos, float kernel_radius, float rho,
    void* stream);

/* The composition antialias.py:106-122 runs for the gradient
 * (compact_valid_ints of prim_ids, dda_compute_fragments,
 * backward_antialiased_cluster_bitset) as one launch sequence with no host
 * synchronization: prim_ids (B*E) and normals (B*E,2) are the outputs of
 * s2c_mark_discontinuity_edges before compaction, and each workgroup DDA-walks
 * the marked edges it holds (dda.cu:9-87, 185-216) straight into the
 * per-sample gradient.  num_views = B, the view count of pos and of the two
 * bitsets.  Same result as the three calls, up to the order of the d_pos
 * atomics; the word-plane path applies as above. */
int s2c_backward_antialiased_cluster_bitset_edges(
    int num_views, int H, int W, int num_slots, const int* bitset,
    const int* target_bitset, int V, const float* pos, int E, const int* edges,
    const int* prim_ids, const float* normals, const int* edge2cluster,
    float* d_pos, float kernel_radius, float rho, void* stream);

/* compact_utils.cu:54-77 (compact_valid_ints), stable order. */
int s2c_compact_valid_ints(int N, const int* input, int* output,
                           int* valid_count_out, void* stream);

/* ---- strands: curve_utils.cu ---------------------------------------------- */

/* curve_utils.cu:135-147 (smooth_strands). Entries whose weight sum is
 * <= 1e-6 are left untouched (the caller zero-fills, curve_utils.py:203). */
int s2c_smooth_strands(int num_strands, int num_points, const float* strands,
                       const float* arclength, float* smoothed_strands,
                       float beta, void* stream);

/* curve_utils.cu:149-166 (skinning). */
int s2c_skinning(int num_strands, int num_points, const float* strands,
                 const float* arclength, const float* handle_strands,
                 const float* canonical_handles, const float* R,
                 float* skinned_strands, float beta, void* stream);

/* ---- fused forms of the reference's call sequences ------------------------
 * No single reference launcher corresponds; each equals the sequence cited,
 * with no fragment list in memory.  None waits for the device. */

/* dda.py:5-76 then rasterize.py:170-214, as used at 3_fitting.py:254-261:
 * all B*E edges DDA-rasterized and OR-ed into bitset (B,H,W,num_slots),
 * which this call clears first.  Equal to the reference's mask on every view
 * (views without fragments are all zero, where the reference shrinks B);
 * cluster ids outside [0, 32*num_slots) set no bit (the reference writes
 * them out of bounds).  Its binning buffers are kept per (device, stream)
 * until s2c_dda_cluster_mask_release; their size follows from the shape,
 * so no call reads anything back.  A workgroup of the walk whose
 * (cell, cluster) keys overflow its table has its fragments OR-ed
 * directly into the bitset by a kernel behind the binned pass (same
 * mask). */
int s2c_dda_cluster_mask(int B, int H, int W, int V, const float* pos, int E,
                         const int* edges, const int* edge2cluster,
                         int num_slots, int* bitset, void* stream);
/* The same with the screen-tile side of the binning fixed to
 * 1 << tile_log2 (2..6; 0 = chosen automatically, as above).  The result
 * does not depend on it; a side whose tile bitset (side^2 * num_slots
 * words) exceeds 128 KB of LDS, or whose (cell, cluster) keys need more
 * than 31 bits, is an S2C_ERR_ARG. */
int s2c_dda_cluster_mask_tiled(int B, int H, int W, int V, const float* pos, int E,
                               const int* edges, const int* edge2cluster, int num_slots,
                               int tile_log2, int* bitset, void* stream);
/* Frees the binning buffers the two calls above keep for `stream` on the
 * current device (after waiting for the stream). */
int s2c_dda_cluster_mask_release(void* stream);

/* rasterize.py:6-53 then rasterize.py:170-214, as used at 3_fitting.py:154-163
 * (render_bitset_from_triangles); clears bitset first. */
int s2c_triangle_cluster_mask(int B, int H, int W, int V, const float* pos,
                              int T, const int* tri, const int* tri2cluster,
                              int num_slots, int* bitset, void* stream);

/* ---- batched / served card fitting -----------------------------------------
 * No reference launcher corresponds: these run, for many independent
 * per-cluster fitting jobs at once, the iteration of 3_fitting.py:381-407
 * (triangle mask, antialias edge gradient, planarity and uniform-Laplacian
 * terms, UniformAdam of optimizer.py:33-68 with ExponentialLR decay), one
 * workgroup per job, with the largesteps reparameterization of
 * 3_fitting.py:347-384 (param = M pos, pos = M^{-1} param, M = I + lam
 * (L_h + w L_w) of the strip) unless lam = 0.
 *
 * Job j: a res[j] x res[j] view (res <= 256), a Q = quads[j] <= 32 strip of
 * 2(Q+1) vertices (row r, column c -> r*(Q+1) + c) stored from
 * verts[vert_off[j]] (x,y,z fp32), a row-major 4x4 clip matrix mvp[16*j],
 * and a packed target mask (bit p = pixel h*res + w of word p/32) starting
 * at target_bits[bits_off[j]].  params (3 per vertex, the largesteps
 * parameter, set from verts at a job's first step), adam_m (3 per vertex),
 * adam_v (1 per vertex) and steps[j] hold the optimizer state; verts holds
 * the positions after each call.  hyper is a HOST array of 10 floats: lr0,
 * beta1, beta2, gamma (lr decay per step), w_pln, w_lap, rho, radius, lam
 * (laplacian_lambda; 0 = optimize positions directly), w_weight
 * (laplacian_w_weight). */

/* Target masks: for job j, num_strands[j] strands of verts_per_strand
 * clip-space float4 vertices from strand_clip[strand_off[j]]; every edge is
 * DDA-rasterized (dda.cu:9-225) into the job's packed mask. */
int s2c_fit_targets(int num_jobs, const int* res, const int* bits_off,
                    const int* strand_off, const int* num_strands,
                    int verts_per_strand, const float* strand_clip,
                    unsigned* target_bits, void* stream);

/* Admits num_jobs jobs into slot tables of `capacity` slots (the tables
 * of s2c_fit_cards; vertex rows of slot s start at s * max_verts).  Job i
 * is records[8i .. 8i+8) = (slot, res, quads, word offset of its target
 * mask, vertex count <= 66, iteration limit, 0, 0) and the 16 + 3 *
 * max_verts floats from frecords[i * (16 + 3 * max_verts)] = its 4 x 4
 * clip matrix (row-major) then its vertices (x, y, z); max_verts >= 66.
 * Writes res, quads, bits_off, mvp, the vertices and limit (skipped when
 * null), zeroes steps and the Adam moments and sets slots[slot] = slot.
 * One launch for a whole admission batch, so a serving loop admits jobs
 * with one host-to-device copy.  No counterpart in the reference: it is
 * the admission step of the serving loop around 3_fitting.py:381-407. */
int s2c_fit_admit(int num_jobs, const int* records, const float* frecords, int capacity,
                  int max_verts, int* slots, int* res, int* quads, int* bits_off, int* limit,
                  float* mvp, int* steps, float* verts, float* adam_m, float* adam_v,
                  void* stream);

/* Runs `iters` fitting iterations for the jobs slots[0..num_slots) (a
 * negative entry is an idle slot and is skipped), then
 * renders the updated card and writes popcount(render XOR target) to
 * loss_out[job]; verts, params, Adam state and steps are updated in place
 * (with iters = 0 only loss_out is written, so measuring a job changes
 * nothing).  With a non-null `limit`, a job runs at most limit[job] -
 * steps[job] of the iterations, and a job with none left is not touched
 * by a call with iters > 0 (its results stay readable while later calls
 * run).  Jobs
 * never share state, and each job's arithmetic order is fixed, so a job's
 * result does not depend on which other jobs run in the same call, on
 * their order in `slots`, or on `chunk`: the silhouette samples a
 * workgroup evaluates per pass (256, 512 or 1024; 0 = 512), a tuning knob
 * trading LDS per workgroup against passes per iteration. */
int s2c_fit_cards(int num_slots, const int* slots, int iters, const int* res,
                  const int* quads, const int* vert_off, const int* bits_off,
                  const float* mvp, float* verts, float* params, float* adam_m,
                  float* adam_v, int* steps, const int* limit, const unsigned* target_bits,
                  const float* hyper, float* loss_out, int chunk, void* stream);

/* compute_bishop_frame (curve_utils.py:104-165, a Python loop over the
 * vertices in the reference): rotation-minimizing frames of num_strands
 * polylines of num_points float3 points by double reflection.  Writes
 * tangents, normals and binormals, each (num_strands, num_points, 3). */
int s2c_bishop_frame(int num_strands, int num_points, const float* strands, float* tangents,
                     float* normals, float* binormals, void* stream);

/* resample_strands_by_arclength (hair_loader.cpp:261-307): num_strands
 * strands of num_samples[i] >= 2 float3 points starting at point
 * offsets[i] (exclusive prefix sum of num_samples) of `points`, resampled
 * to target_num_samples points evenly spaced in arclength into
 * out (num_strands, target_num_samples, 3).  `lengths` is caller-provided
 * scratch of sum(num_samples) floats. */
int s2c_resample_strands_by_arclength(int num_strands, const int* num_samples,
                                      const long long* offsets, const float* points,
                                      int target_num_samples, float* lengths, float* out,
                                      void* stream);

/* ---------------------------------------------------------------------- */
/* Differentiable card coverage (hand-written adjoint)                     */
/* ---------------------------------------------------------------------- */
/* No single reference entry point: the differentiable counterpart of the
 * rasterize (rasterize.py:6-95) + antialias (antialias.py:66-121) chain
 * that 3_fitting.py:380-400 differentiates.  Coverage is area-sampled:
 * coverage[b, h, w] += area(triangle ∩ [w, w+1] x [h, h+1]) for every
 * triangle of view b, in the screen coordinates of dda.cu:33-36.  It is
 * continuous in the vertices, so its gradient is checkable by finite
 * differences.  pos: (B, V, 4) clip space (w > 0; other triangles are
 * skipped), tri: (T, 3) indices in [0, V), H * W <= 8,388,607.  Both
 * calls ACCUMULATE (coverage and d_pos must be zero-filled by the caller
 * for a fresh result).  The backward recomputes all geometry from pos and
 * writes d loss / d pos (x, y, w; z gets 0) given d_coverage = d loss /
 * d coverage. */
int s2c_card_coverage_forward(int B, int H, int W, int V, const float* pos, int T,
                              const int* tri, float* coverage, void* stream);
int s2c_card_coverage_backward(int B, int H, int W, int V, const float* pos, int T,
                               const int* tri, const float* d_coverage, float* d_pos,
                               void* stream);

/* ---------------------------------------------------------------------- */
/* Large-steps parameterization of grid meshes                            */
/* ---------------------------------------------------------------------- */
/* M = I + lam * (L_h + w_weight * L_w) of 3_fitting.py:347-356, for the
 * concatenated grid meshes of texture_utils.py:525-655: grid g has rows[g]
 * rows of cols[g] vertices, vertex (i, j) at voff[g] + i * cols[g] + j;
 * L_h / L_w are the uniform Laplacians (optimizer.py:70-99) of its h_edges
 * (i, j)-(i, j+1) and w_edges (i, j)-(i+1, j).  x and u are (V, 3) float32
 * with V = voff[last] + rows[last] * cols[last]; rows, cols, voff and
 * fac_off are device int32 arrays of num_grids entries.  lam, w_weight >= 0.
 *
 * s2c_grid_system_apply: u = M x.  Replaces largesteps.parameterize's
 * to_differential(M, x) (called at 3_fitting.py:365).
 *
 * s2c_grid_system_factor: the factor of M for the solve, computed once per
 * (lam, w_weight) and layout, as the reference builds its Cholesky solver
 * once per M.  Writes rows[g] * cols[g] floats at factor + fac_off[g] (the
 * pivot reciprocals of the grid's decoupled tridiagonal systems); they
 * depend only on (rows[g], cols[g], lam, w_weight), so grids of one shape
 * may share one offset (and the call may list each shape once).
 *
 * s2c_grid_system_solve: x = M^{-1} u with the factor of the same lam.
 * Replaces largesteps.parameterize's from_differential(M, u)
 * (3_fitting.py:384, 457).  max_verts / max_rows bound rows[g] * cols[g]
 * and rows[g] over all grids and size the workgroup's LDS
 * (24 * max_verts + 4 * max_rows^2 bytes, at most 65536: e.g. 2 rows of up
 * to 1365 columns); a grid beyond them is written as NaN. */
int s2c_grid_system_apply(int num_grids, const int* rows, const int* cols, const int* voff,
                          float lam, float w_weight, const float* x, float* u, void* stream);
/* y = (L_h + w_weight L_w) x, the grids' uniform Laplacians (no identity,
 * no lam): with w_weight = 1 the laplacian_edges of h_edges + w_edges that
 * 3_fitting.py:359 builds for its w_lap term, applied without a sparse
 * matrix.  Symmetric, so it is also its own adjoint. */
int s2c_grid_laplacian(int num_grids, const int* rows, const int* cols, const int* voff,
                       float w_weight, const float* x, float* y, void* stream);
/* D (Q, 3) = (p0 + p2) - (p1 + p3) for every quad [p0, p1, p2, p3] =
 * [(i,j), (i+1,j), (i+1,j+1), (i,j+1)] of the grids (the planarity term of
 * 3_fitting.py:181-199 over the quads of texture_utils.py:557-565); grid
 * g's quads are rows qoff[g] .. qoff[g] + (rows-1)(cols-1) in row-major
 * (i, j) order.  The backward writes dpos (V, 3) = the adjoint of D for
 * dD, gathered per vertex (no atomics). */
int s2c_grid_quad_diagonals(int num_grids, const int* rows, const int* cols, const int* voff,
                            const int* qoff, const float* pos, float* D, void* stream);
int s2c_grid_quad_diagonals_backward(int num_grids, const int* rows, const int* cols,
                                     const int* voff, const int* qoff, const float* dD,
                                     float* dpos, void* stream);

/* optimizer.py:33-68 (UniformAdam.step) for one parameter of n points x dim
 * channels, as a device step with no host synchronization: g1 (n, dim) and
 * g2 (n) are its state, updated in place; `step` is the step count after the
 * increment (>= 1); lr, b1, b2 are the optimizer group's values as the
 * reference's Python floats, rounded to float32 once (b1 ** step included).
 * param <- param - lr * (g1 / (1 - b1^step)) / (max_i sqrt(g2_i /
 * (1 - b2^step)) + 1e-8); a NaN in any g2 makes every update NaN, as
 * torch's max() does. */
int s2c_uniform_adam_step(int n, int dim, double lr, double b1, double b2,
                          int step, float* param, const float* grad,
                          float* g1, float* g2, void* stream);
int s2c_grid_system_factor(int num_grids, const int* rows, const int* cols, const int* fac_off,
                           float lam, float w_weight, float* factor, void* stream);
int s2c_grid_system_solve(int num_grids, const int* rows, const int* cols, const int* voff,
                          const int* fac_off, const float* factor, int max_verts, int max_rows,
                          float lam, const float* u, float* x, void* stream);

/* ---------------------------------------------------------------------- */
/* Card mesh inputs                                                       */
/* ---------------------------------------------------------------------- */
/* Replaces strand_to_strip_mesh (2_texture.py:17-96, host numpy in the
 * reference) for many strands: strand s has counts[s] >= 2 float32 xyz
 * points at points + 3 * offsets[s] (int64 point offsets, the prefix of
 * counts).  For every strand and each of the num_rot (1..16) tangent
 * rotations rot_deg (a HOST array, degrees) one (res_w, counts[s], 3)
 * strip is written at out + 3 * res_w * (num_rot * offsets[s] +
 * r * counts[s]): the strips in the order 2_texture.py:304-312 collects
 * them, i.e. the vertex array of their combined grid meshes.  res_w >= 2;
 * width as the reference's float argument. */
int s2c_strands_to_strips(int num_strands, const int* counts, const long long* offsets,
                          const float* points, int res_w, double width, int num_rot,
                          const float* rot_deg, float* out, void* stream);

/* Replaces wisp_to_grid_mesh + combine_wisps_with_cluster_labels
 * (texture_utils.py:525-655, host numpy): grid g has rows[g] x cols[g]
 * vertices from voff[g]; its (rows-1)(cols-1) quads start at qoff[g]
 * (faces at 2 * qoff[g]), its rows*(cols-1) h_edges at hoff[g] and its
 * (rows-1)*cols w_edges at woff[g] (int32 arrays, prefixes of those
 * counts).  faces (F,3), quads (Q,4), h_edges (H,2), w_edges (W,2) int32.
 * Optional pairs (both NULL or both given): grid2cluster (int64 per grid)
 * -> face2cluster (int64 per face); rects (float64 [u0, v0, u1, v1] per
 * grid) -> uvs (float64 (V,2)), with v1 reduced by margin_top. */
int s2c_grid_mesh_topology(int num_grids, const int* rows, const int* cols, const int* voff,
                           const int* qoff, const int* hoff, const int* woff,
                           const long long* grid2cluster, const double* rects, double margin_top,
                           int* faces, int* quads, int* h_edges, int* w_edges,
                           long long* face2cluster, double* uvs, void* stream);

/* Replaces build_clustered_strand_edges (3_fitting.py:17-60): edges (E,2)
 * between consecutive points of each strand (counts[s] >= 1 points from
 * point offset offsets[s]) and edge2cluster[e] = cluster_ids[strand]. */
int s2c_strand_edges(int num_strands, const int* counts, const long long* offsets,
                     const int* cluster_ids, int* edges, int* edge2cluster, void* stream);

/* ---------------------------------------------------------------------- */
/* Software-MSAA texture patches (render_texture, texture_utils.py:106-288) */
/* ---------------------------------------------------------------------- */
/* The reference rasterizes each wisp's tube mesh once per sample of an
 * 8 x 8 jitter grid (64 rasterize + interpolate calls) and accumulates with
 * torch.  Here: s2c_msaa_jitter makes the jittered copies (num_samples, V, 4)
 * = (x + dx[s], y + dy[s], z, 1) of the (V, 3) NDC vertices (dx, dy:
 * device float32 arrays) for one batched rasterize; s2c_msaa_resolve
 * reduces the (num_samples, H, W, 4) raster over its samples per pixel, in
 * sample order, into rgb (H, W, 3) = 0.5 (mean interpolated attr over the
 * covered samples + 1) ((0, 1, 0) + 1 halved where none) and alpha (H, W)
 * = sum of weights[s] over covered samples (attr (V, 3), tri (T, 3),
 * weights device float32); s2c_blur3x3_reflect blurs with the separable
 * 3-tap kernel3 (a HOST array), reflect borders, x then y (tmp, out:
 * (H, W)); s2c_atlas_paste writes the (atlas_y1 - atlas_y0) x (atlas_x1 -
 * atlas_x0) rectangle of the R x R byte atlases from the vertically
 * flipped patch starting at (crop_y0, crop_x0), values clamped to [0, 1]
 * and scaled by 255 with truncation. */
int s2c_msaa_jitter(int V, const float* pos, int num_samples, const float* dx, const float* dy,
                    float* out, void* stream);
int s2c_msaa_resolve(int num_samples, int H, int W, const float* rast, const float* attr,
                     const int* tri, const float* weights, float* rgb, float* alpha,
                     void* stream);
int s2c_blur3x3_reflect(int H, int W, const float* kernel3, const float* in, float* tmp,
                        float* out, void* stream);
/* The wisp masks of render_patches_reduced (texture_utils.py:290-411) for
 * num_wisps (<= 65535) wisps in one call: wisp g has rows[g] strands of
 * cols[g] points, clip-space float4 vertices from pos + 4 * voff[g] (int64
 * vertex offsets), and the edges of build_edge_indices_from_wisp
 * (__init__.py:165-175).  out (num_wisps, H, W) bytes = the DDA fragments'
 * depth test and the interpolation of a constant 1 (dda.cu), flipped
 * vertically, clamped and scaled by 255 with truncation, 0 where nothing
 * is drawn.  keys: (num_wisps, H, W) int64 scratch; max_edges bounds
 * rows[g] * (cols[g] - 1). */
int s2c_patch_masks(int num_wisps, int H, int W, const int* rows, const int* cols,
                    const long long* voff, int max_edges, const float* pos, long long* keys,
                    unsigned char* out, void* stream);
/* The fused form of rasterize + s2c_msaa_resolve for the jittered views
 * (num_samples, V, 4) of s2c_msaa_jitter: no fragment lists; per (view,
 * pixel) the minimum (bits(z + 2) << 32 | triangle) over the covering
 * triangles, which picks the fragment the drop-in depth test keeps
 * (rasterize.cu:198-228; at one pixel its fragments of one view are ordered
 * by triangle), in keys (num_samples, H, W) int64 scratch; then the resolve
 * with the winners' barycentrics recomputed.  Same rgb / alpha as
 * rasterize + s2c_msaa_resolve. */
int s2c_msaa_patch(int V, int T, const float* jittered, const int* tri, const float* attr,
                   int num_samples, int H, int W, const float* weights, long long* keys,
                   float* rgb, float* alpha, void* stream);
int s2c_atlas_paste(int H, int W, const float* rgb, const float* alpha, int crop_x0, int crop_y0,
                    int atlas_x0, int atlas_y0, int atlas_x1, int atlas_y1, int R,
                    unsigned char* atlas_rgb, unsigned char* atlas_alpha, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* STRNS2CARDS_HIP_H */
