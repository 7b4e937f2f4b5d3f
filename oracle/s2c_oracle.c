/*
 * CPU oracle for the pystrns2cards hot path — TEST INFRASTRUCTURE ONLY.
 *
 * This file restates, element for element, the arithmetic of the reference
 * CUDA kernels in kenji-tojo/strns2cards (paths relative to that repository).
 * The tests compare the HIP library against it, and bench.py times it as
 * the CPU baseline; nothing in the shipped library links it.
 *
 * Evaluation policy (shared with the HIP kernels, which is what makes the
 * integer outputs bit-exact): every float expression is evaluated in the
 * order written in the reference, one IEEE-754 binary32 operation at a time
 * (built with -ffp-contract=off, no fast-math), float->int conversions
 * saturate (NaN -> 0), and every list output is in the canonical stable
 * order (primitive-major, then candidate index).  The reference's own
 * compaction order is nondeterministic across thread blocks
 * (rasterize.cu:315-318, compact_utils.cu:39-42); the stable order is one of
 * its possible outputs.
 *
 * Parity pins: see tests/test_oracle_*.py (closed forms, brute force,
 * invariants).  Functions whose only pin is a tolerance are stated there.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <limits.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#define S2C_MIN(a, b) ((a) < (b) ? (a) : (b))
#define S2C_MAX(a, b) ((a) > (b) ? (a) : (b))

/* Saturating float->int32 truncation with NaN -> 0: the conversion that
 * cvt.rzi.s32.f32 (the reference's static_cast<int> on its GPU) and
 * v_cvt_i32_f32 both perform. */
static inline int f2i(float x) {
    if (x != x) return 0;
    if (x >= 2147483648.0f) return INT_MAX;
    if (x <= -2147483648.0f) return INT_MIN;
    return (int)x;
}

/* ------------------------------------------------------------------------ */
/* DDA line rasterization — dda.cu                                           */
/* ------------------------------------------------------------------------ */

/* dda.cu:9-87 (dda_compute_span_kernel) for one primitive.  Returns the
 * candidate fragment count; writes slopes/spans except for zero-length
 * edges, where the reference writes nothing. */
static int dda_span_one(int prim_id, int H, int W, int V, const float *pos,
                        int E, const int *edges, float *slope2, float *span4) {
    const int b = prim_id / E;
    const int e = prim_id % E;
    const int v0 = edges[e * 2 + 0];
    const int v1 = edges[e * 2 + 1];
    const float *p = pos + (size_t)b * V * 4;
    const float x0n = p[v0 * 4 + 0] / p[v0 * 4 + 3];
    const float y0n = p[v0 * 4 + 1] / p[v0 * 4 + 3];
    const float x1n = p[v1 * 4 + 0] / p[v1 * 4 + 3];
    const float y1n = p[v1 * 4 + 1] / p[v1 * 4 + 3];
    const float fH = (float)H, fW = (float)W;
    float x0 = 0.5f * (x0n + 1.f) * fW;
    float y0 = 0.5f * (1.f - y0n) * fH;
    float x1 = 0.5f * (x1n + 1.f) * fW;
    float y1 = 0.5f * (1.f - y1n) * fH;
    const float dx = x1 - x0;
    const float dy = y1 - y0;
    if (dx == 0.f && dy == 0.f) return 0;
    if (fabsf(dx) >= fabsf(dy)) {
        const float s = dy / dx;
        slope2[0] = 0.f;
        slope2[1] = s;
        if (x0 > x1) { float t = x0; x0 = x1; x1 = t; t = y0; y0 = y1; y1 = t; }
        const float sx = S2C_MAX(floorf(x0), 0.f);
        const float sy = y0 + s * (sx - x0);
        const int i0 = f2i(sx);
        const int i1 = S2C_MIN(f2i(floorf(x1)), W - 1) + 1;
        span4[0] = sx; span4[1] = sy; span4[2] = x0 - sx; span4[3] = x1 - sx;
        return S2C_MAX(i1 - i0, 0);
    } else {
        const float s = dx / dy;
        slope2[0] = 1.f;
        slope2[1] = s;
        if (y0 > y1) { float t = x0; x0 = x1; x1 = t; t = y0; y0 = y1; y1 = t; }
        const float sy = S2C_MAX(floorf(y0), 0.f);
        const float sx = x0 + s * (sy - y0);
        const int i0 = f2i(sy);
        const int i1 = S2C_MIN(f2i(floorf(y1)), H - 1) + 1;
        span4[0] = sx; span4[1] = sy; span4[2] = y0 - sy; span4[3] = y1 - sy;
        return S2C_MAX(i1 - i0, 0);
    }
}

/* dda.cu:89-122.  Inclusive prefix of the counts; returns the total
 * (0 for zero primitives, where the reference reads index -1). */
int orc_dda_compute_span(int P, const int *edge_ids, int H, int W, int V,
                         const float *pos, int E, const int *edges,
                         int *prefix, float *slopes, float *spans) {
    long long acc = 0;
    for (int i = 0; i < P; ++i) {
        acc += dda_span_one(edge_ids[i], H, W, V, pos, E, edges,
                            slopes + (size_t)i * 2, spans + (size_t)i * 4);
        prefix[i] = (int)acc;
    }
    return P > 0 ? prefix[P - 1] : 0;
}

/* dda.cu:124-144 (perspective_correct_interpolate). */
static float dda_t1(int H, int W, const float *p0, const float *p1,
                    float x, float y, float axis) {
    float t0, t1;
    if (axis == 0.f) {
        const float qx = -1.f + 2.f * (x / (float)W);
        const float a = p0[0] - qx * p0[3];
        const float c = p1[0] - qx * p1[3];
        t0 = fabsf(c);
        t1 = fabsf(a);
    } else {
        const float qy = -1.f + 2.f * (y / (float)H);
        const float a = -p0[1] - qy * p0[3];
        const float c = -p1[1] - qy * p1[3];
        t0 = fabsf(c);
        t1 = fabsf(a);
    }
    const float denom = t0 + t1;
    if (denom < 1e-6f) return 0.5f;
    return t1 / denom;
}

/* dda.cu:146-225 for fragment `k` of a primitive with known span.  Returns 1
 * and fills the outputs when the fragment survives the screen and z tests. */
static int dda_frag_one(int k, int b, int e, const float *slope2,
                        const float *span4, int H, int W, int V,
                        const float *pos, const int *edges,
                        int *pix3, float *attr4) {
    const float d = fminf(fmaxf((float)k + 0.5f, span4[2]), span4[3]);
    const float axis = slope2[0];
    const float s = slope2[1];
    const float ddx = (axis == 0.f) ? d : d * s;
    const float ddy = (axis == 0.f) ? d * s : d;
    const float x = span4[0] + ddx;
    const float y = span4[1] + ddy;
    const int h = f2i(floorf(y));
    const int w = f2i(floorf(x));
    if (h < 0 || h >= H || w < 0 || w >= W) return 0;
    const int v0 = edges[e * 2 + 0];
    const int v1 = edges[e * 2 + 1];
    const float *pp = pos + (size_t)b * V * 4;
    float p0[4], p1[4];
    for (int i = 0; i < 4; ++i) { p0[i] = pp[v0 * 4 + i]; p1[i] = pp[v1 * 4 + i]; }
    const float t1 = dda_t1(H, W, p0, p1, x, y, axis);
    const float t0 = 1.f - t1;
    const float pz = t0 * p0[2] + t1 * p1[2];
    const float pw = t0 * p0[3] + t1 * p1[3];
    const float zw = pz / pw;
    if (!isfinite(zw) || zw < -1.f || zw > 1.f) return 0;
    pix3[0] = b; pix3[1] = h; pix3[2] = w;
    attr4[0] = axis; attr4[1] = t1; attr4[2] = zw; attr4[3] = (float)(e + 1);
    return 1;
}

/* dda.cu:146-256.  Candidates that fail a test keep the caller's -1 fill in
 * frag_pix (dda.py:60) and leave frag_attrs untouched. */
void orc_dda_compute_fragments(int P, int N, const int *prefix,
                               const int *edge_ids, const float *slopes,
                               const float *spans, int H, int W, int V,
                               const float *pos, int E, const int *edges,
                               int *frag_pix, float *frag_attrs) {
    int base = 0;
    for (int p = 0; p < P && base < N; ++p) {
        const int end = prefix[p];
        const int b = edge_ids[p] / E, e = edge_ids[p] % E;
        for (int f = base; f < end && f < N; ++f)
            dda_frag_one(f - base, b, e, slopes + (size_t)p * 2,
                         spans + (size_t)p * 4, H, W, V, pos, edges,
                         frag_pix + (size_t)f * 3, frag_attrs + (size_t)f * 4);
        base = end;
    }
}

/* dda.cu:258-284. */
void orc_dda_interpolate_attributes(int B, int H, int W, int C,
                                    const float *rast, const float *attr,
                                    const int *edges, float *image) {
    const long long n = (long long)B * H * W;
    for (long long i = 0; i < n; ++i) {
        const float t1 = rast[i * 4 + 1];
        const float t0 = 1.f - t1;
        const int e = f2i(rast[i * 4 + 3]) - 1;
        if (e < 0) continue;
        const float *a0 = attr + (size_t)edges[e * 2 + 0] * C;
        const float *a1 = attr + (size_t)edges[e * 2 + 1] * C;
        for (int c = 0; c < C; ++c) image[i * C + c] = t0 * a0[c] + t1 * a1[c];
    }
}

/* dda.cu:300-327 (serial accumulation in pixel order).  `absum` (may be
 * NULL) receives each d_attr element's sum of |term|, the scale of the
 * rounding of any other summation order (the tests' elementwise bound). */
void orc_backward_dda_interpolate_attributes(int B, int H, int W,
                                             const float *rast, int C,
                                             const float *d_image,
                                             const int *edges, float *d_attr,
                                             float *absum) {
    const long long n = (long long)B * H * W;
    for (long long i = 0; i < n; ++i) {
        const float t1 = rast[i * 4 + 1];
        const float t0 = 1.f - t1;
        const int e = f2i(rast[i * 4 + 3]) - 1;
        if (e < 0) continue;
        float *d0 = d_attr + (size_t)edges[e * 2 + 0] * C;
        float *d1 = d_attr + (size_t)edges[e * 2 + 1] * C;
        for (int c = 0; c < C; ++c) {
            const float g0 = d_image[i * C + c] * t0, g1 = d_image[i * C + c] * t1;
            d0[c] += g0;
            d1[c] += g1;
            if (absum) {
                absum[(size_t)edges[e * 2 + 0] * C + c] += fabsf(g0);
                absum[(size_t)edges[e * 2 + 1] * C + c] += fabsf(g1);
            }
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Triangle rasterization — rasterize.cu                                     */
/* ------------------------------------------------------------------------ */

/* rasterize.cu:9-59 for one (b, t). */
static int tri_rect_one(int b, int t, int H, int W, int V, const float *pos,
                        const int *tri, int *rect4) {
    const float fH = (float)H, fW = (float)W;
    float x0 = fW, y0 = fH, x1 = 0.f, y1 = 0.f;
    for (int i = 0; i < 3; ++i) {
        const float *p = pos + ((size_t)b * V + tri[t * 3 + i]) * 4;
        const float inv_w = 1.f / p[3];
        const float x = 0.5f * (1.f + p[0] * inv_w) * fW;
        const float y = 0.5f * (1.f - p[1] * inv_w) * fH;
        x0 = isnan(x) ? x0 : fminf(x, x0);
        x1 = isnan(x) ? x1 : fmaxf(x, x1);
        y0 = isnan(y) ? y0 : fminf(y, y0);
        y1 = isnan(y) ? y1 : fmaxf(y, y1);
    }
    const int h0 = S2C_MAX(f2i(floorf(y0)), 0);
    const int h1 = S2C_MIN(f2i(floorf(y1)), H - 1) + 1;
    const int w0 = S2C_MAX(f2i(floorf(x0)), 0);
    const int w1 = S2C_MIN(f2i(floorf(x1)), W - 1) + 1;
    const int hl = S2C_MAX(h1 - h0, 0);
    const int wl = S2C_MAX(w1 - w0, 0);
    rect4[0] = h0; rect4[1] = hl; rect4[2] = w0; rect4[3] = wl;
    return hl * wl;
}

/* rasterize.cu:61-88. */
int orc_compute_triangle_rects(int H, int W, int B, int V, const float *pos,
                               int T, const int *tri, int *rects, int *prefix) {
    long long acc = 0;
    for (int i = 0; i < B * T; ++i) {
        acc += tri_rect_one(i / T, i % T, H, W, V, pos, tri, rects + (size_t)i * 4);
        prefix[i] = (int)acc;
    }
    return B * T > 0 ? prefix[B * T - 1] : 0;
}

/* rasterize.cu:132-171 for one candidate pixel (x, y) of triangle (b, t). */
static int tri_frag_one(int b, int t, int x, int y, int H, int W, int V,
                        const float *pos, const int *tri, int *pix3,
                        float *attr4) {
    if (x < 0 || x >= W || y < 0 || y >= H) return 0;
    const float px = -1.f + 2.f * ((float)x + 0.5f) / (float)W;
    const float py = -1.f + 2.f * ((float)y + 0.5f) / (float)H;
    const float *bp = pos + (size_t)b * V * 4;
    const float *P0 = bp + (size_t)tri[t * 3 + 0] * 4;
    const float *P1 = bp + (size_t)tri[t * 3 + 1] * 4;
    const float *P2 = bp + (size_t)tri[t * 3 + 2] * 4;
    const float q0x = P0[0] - px * P0[3], q0y = -P0[1] - py * P0[3];
    const float q1x = P1[0] - px * P1[3], q1y = -P1[1] - py * P1[3];
    const float q2x = P2[0] - px * P2[3], q2y = -P2[1] - py * P2[3];
    const float A0 = q1x * q2y - q1y * q2x;
    const float A1 = q2x * q0y - q2y * q0x;
    const float A2 = q0x * q1y - q0y * q1x;
    const float A = A0 + A1 + A2;
    if (A == 0.f) return 0;
    const float invA = 1.f / A;
    const float b0 = A0 * invA;
    const float b1 = A1 * invA;
    const float zc = A0 * P0[2] + A1 * P1[2] + A2 * P2[2];
    const float wc = A0 * P0[3] + A1 * P1[3] + A2 * P2[3];
    const float z = zc / wc;
    if (!isfinite(b0) || !isfinite(b1) || !isfinite(z)) return 0;
    if (b0 < 0.f || b1 < 0.f || b0 + b1 > 1.f || z < -1.f || z > 1.f) return 0;
    pix3[0] = b; pix3[1] = y; pix3[2] = x;
    attr4[0] = b0; attr4[1] = b1; attr4[2] = z; attr4[3] = (float)(t + 1);
    return 1;
}

/* rasterize.cu:90-172. */
void orc_compute_fragments(int H, int W, int V, const float *pos, int T,
                           const int *tri, int num_tris, int N,
                           const int *prefix, const int *rects,
                           int *frag_pix, float *frag_attrs) {
    int base = 0;
    for (int i = 0; i < num_tris && base < N; ++i) {
        const int end = prefix[i];
        const int *r = rects + (size_t)i * 4;
        if (r[1] > 0 && r[3] > 0) {
            for (int f = base; f < end && f < N; ++f) {
                const int l = f - base;
                tri_frag_one(i / T, i % T, r[2] + l % r[3], r[0] + l / r[3], H, W,
                             V, pos, tri, frag_pix + (size_t)f * 3,
                             frag_attrs + (size_t)f * 4);
            }
        }
        base = end;
    }
}

/* rasterize.cu:197-278.  rast must be zero-filled by the caller
 * (rasterize.py:92).  Key = (bits(z + 2) << 32) | fragment index; the
 * smallest signed key wins. */
void orc_depth_test(int B, int H, int W, int N, const int *frag_pix,
                    const float *frag_attrs, float *rast) {
    const long long n = (long long)B * H * W;
    long long *key = (long long *)malloc(sizeof(long long) * (size_t)(n > 0 ? n : 1));
    for (long long i = 0; i < n; ++i) key[i] = LLONG_MAX;
    for (int f = 0; f < N; ++f) {
        const float zz = frag_attrs[(size_t)f * 4 + 2] + 2.f;
        int32_t zb;
        memcpy(&zb, &zz, 4);
        const long long k = (long long)(((uint64_t)(uint32_t)zb << 32) | (uint64_t)(uint32_t)f);
        const long long pix = ((long long)frag_pix[f * 3 + 0] * H + frag_pix[f * 3 + 1]) * W
                              + frag_pix[f * 3 + 2];
        if (k < key[pix]) key[pix] = k;
    }
    for (long long i = 0; i < n; ++i) {
        const int f = (int)(uint32_t)(key[i] & 0xFFFFFFFFLL);
        if (f < 0) continue;
        memcpy(rast + i * 4, frag_attrs + (size_t)f * 4, 16);
    }
    free(key);
}

/* rasterize.cu:280-365, stable order. */
int orc_filter_valid_fragments(int N, const int *fp, const float *fa,
                               int *fp_out, float *fa_out) {
    int o = 0;
    for (int i = 0; i < N; ++i) {
        if (fp[(size_t)i * 3] < 0) continue;
        memcpy(fp_out + (size_t)o * 3, fp + (size_t)i * 3, 12);
        memcpy(fa_out + (size_t)o * 4, fa + (size_t)i * 4, 16);
        ++o;
    }
    return o;
}

/* rasterize.cu:367-400. */
void orc_interpolate_triangle_attributes(int B, int H, int W, int C,
                                         const float *rast, const float *attr,
                                         const int *tri, float *image) {
    const long long n = (long long)B * H * W;
    for (long long i = 0; i < n; ++i) {
        const float b0 = rast[i * 4 + 0], b1 = rast[i * 4 + 1];
        const float b2 = 1.f - b0 - b1;
        const int t = f2i(rast[i * 4 + 3]) - 1;
        if (t < 0) continue;
        const float *a0 = attr + (size_t)tri[t * 3 + 0] * C;
        const float *a1 = attr + (size_t)tri[t * 3 + 1] * C;
        const float *a2 = attr + (size_t)tri[t * 3 + 2] * C;
        for (int c = 0; c < C; ++c)
            image[i * C + c] = b0 * a0[c] + b1 * a1[c] + b2 * a2[c];
    }
}

/* rasterize.cu:402-438 (serial accumulation in pixel order); `absum` as in
 * orc_backward_dda_interpolate_attributes. */
void orc_backward_interpolate_triangle_attributes(int B, int H, int W, int C,
                                                  const float *rast,
                                                  const float *d_image,
                                                  const int *tri, float *d_attr,
                                                  float *absum) {
    const long long n = (long long)B * H * W;
    for (long long i = 0; i < n; ++i) {
        const float b0 = rast[i * 4 + 0], b1 = rast[i * 4 + 1];
        const float b2 = 1.f - b0 - b1;
        const int t = f2i(rast[i * 4 + 3]) - 1;
        if (t < 0) continue;
        float *g0 = d_attr + (size_t)tri[t * 3 + 0] * C;
        float *g1 = d_attr + (size_t)tri[t * 3 + 1] * C;
        float *g2 = d_attr + (size_t)tri[t * 3 + 2] * C;
        for (int c = 0; c < C; ++c) {
            const float g = d_image[i * C + c];
            g0[c] += g * b0; g1[c] += g * b1; g2[c] += g * b2;
            if (absum) {
                absum[(size_t)tri[t * 3 + 0] * C + c] += fabsf(g * b0);
                absum[(size_t)tri[t * 3 + 1] * C + c] += fabsf(g * b1);
                absum[(size_t)tri[t * 3 + 2] * C + c] += fabsf(g * b2);
            }
        }
    }
}

/* rasterize.cu:474-520.  bitset must be zero-filled by the caller. */
void orc_cluster_mask_from_fragments(int N, const int *fp, const float *fa,
                                     const int *prim2cluster, int H, int W,
                                     int S, int *bitset) {
    for (int f = 0; f < N; ++f) {
        const int prim = f2i(fa[(size_t)f * 4 + 3]) - 1;
        if (prim < 0) continue;
        const int k = prim2cluster[prim];
        const long long idx = (((long long)fp[f * 3 + 0] * H + fp[f * 3 + 1]) * W
                               + fp[f * 3 + 2]) * S + k / 32;
        bitset[idx] |= (int)(1u << (k % 32));
    }
}

/* ------------------------------------------------------------------------ */
/* Bitset utilities — bitset_utils.cu                                        */
/* ------------------------------------------------------------------------ */

/* bitset_utils.cu:6-35 (colors summed in ascending cluster id). */
void orc_accumulate_bitset_rgb(int B, int H, int W, int S, const int *bitset,
                               const float *rgb, float *out) {
    const long long n = (long long)B * H * W;
    for (long long i = 0; i < n; ++i) {
        float r = 0.f, g = 0.f, bl = 0.f;
        for (int s = 0; s < S; ++s) {
            const uint32_t word = (uint32_t)bitset[i * S + s];
            if (!word) continue;
            for (int j = 0; j < 32; ++j) {
                if (word & (1u << j)) {
                    const int k = s * 32 + j;
                    r += rgb[k * 3 + 0]; g += rgb[k * 3 + 1]; bl += rgb[k * 3 + 2];
                }
            }
        }
        out[i * 3 + 0] = r; out[i * 3 + 1] = g; out[i * 3 + 2] = bl;
    }
}

/* bitset_utils.cu:37-52. */
void orc_popcount_bitset(int B, int H, int W, int S, const int *bitset,
                         float *out) {
    const long long n = (long long)B * H * W;
    for (long long i = 0; i < n; ++i) {
        int c = 0;
        for (int s = 0; s < S; ++s) c += __builtin_popcount((uint32_t)bitset[i * S + s]);
        out[i] = (float)c;
    }
}

/* ------------------------------------------------------------------------ */
/* Antialias — antialias.cu, compact_utils.cu                                */
/* ------------------------------------------------------------------------ */

static int third_vertex(const int *tri, int t, int v0, int v1) {
    const int i0 = tri[t * 3 + 0], i1 = tri[t * 3 + 1], i2 = tri[t * 3 + 2];
    return (i0 != v0 && i0 != v1) ? i0 : (i1 != v0 && i1 != v1) ? i1
         : (i2 != v0 && i2 != v1) ? i2 : -1;
}

/* antialias.cu:29-35. */
static void project(const float *p, int v, float fW, float fH, float *x, float *y) {
    const float w = p[v * 4 + 3];
    const float nx = p[v * 4 + 0] / w;
    const float ny = p[v * 4 + 1] / w;
    *x = fW * 0.5f * (1.f + nx);
    *y = fH * 0.5f * (1.f - ny);
}

/* antialias.cu:38-119.  prim_ids must be -1-filled and normals zero-filled
 * by the caller (antialias.py:36-37). */
void orc_mark_discontinuity_edges(int H, int W, int B, int V, const float *pos,
                                  const int *tri, int E, const int *edges,
                                  const int *edge2tri, int *prim_ids,
                                  float *normals) {
    const float fH = (float)H, fW = (float)W;
    for (int idx = 0; idx < B * E; ++idx) {
        const int b = idx / E, e = idx % E;
        const float *p = pos + (size_t)b * V * 4;
        const int v0 = edges[e * 2 + 0], v1 = edges[e * 2 + 1];
        float x0, y0, x1, y1;
        project(p, v0, fW, fH, &x0, &y0);
        project(p, v1, fW, fH, &x1, &y1);
        const float ex = x1 - x0, ey = y1 - y0;
        const float len = sqrtf(ex * ex + ey * ey);
        if (len == 0.f) continue;
        const float nx = -ey / len;
        const float ny = ex / len;
        float d0 = 0.f, d1 = 0.f;
        const int t0 = edge2tri[e * 2 + 0], t1 = edge2tri[e * 2 + 1];
        if (t0 >= 0) {
            float x2, y2;
            project(p, third_vertex(tri, t0, v0, v1), fW, fH, &x2, &y2);
            d0 = nx * (x2 - x0) + ny * (y2 - y0);
        }
        if (t1 >= 0) {
            float x2, y2;
            project(p, third_vertex(tri, t1, v0, v1), fW, fH, &x2, &y2);
            d1 = nx * (x2 - x0) + ny * (y2 - y0);
        }
        if (d0 * d1 >= 0.f) prim_ids[idx] = idx;
        const float sign = 2.f * (.5f - (float)(d0 > 0.f || d1 > 0.f));
        normals[(size_t)idx * 2 + 0] = sign * nx;
        normals[(size_t)idx * 2 + 1] = sign * ny;
    }
}

/* compact_utils.cu:7-77, stable order. */
int orc_compact_valid_ints(int N, const int *in, int *out) {
    int o = 0;
    for (int i = 0; i < N; ++i)
        if (in[i] >= 0) out[o++] = in[i];
    return o;
}

/* antialias.cu:121-132 (libm expf in place of the fast exponential). */
static float gauss_w(float x, float y, int w, int h, float R) {
    const float pi = 3.14159265358979323846f;
    const float sigma = R / 4.f;
    const float c = 1.f / (sigma * sigma * 2.f * pi);
    const float alpha = 0.5f / (sigma * sigma);
    const float dx = (float)w + 0.5f - x;
    const float dy = (float)h + 0.5f - y;
    const float n2 = dx * dx + dy * dy;
    return c * expf(-alpha * n2) * (float)(n2 < R * R);
}

static float bit_of(const int *bs, int b, int w, int h, int k, int H, int W, int S) {
    const long long pix = (long long)H * W * b + w + (long long)h * W;
    return (float)(((uint32_t)bs[pix * S + k / 32] >> (k % 32)) & 1u);
}

/* antialias.cu:147-176. */
static float adjoint_sum(int b, float x, float y, int k, int H, int W, int S,
                         const int *bs, const int *tg, float R, float rho) {
    const int Ri = f2i(floorf(R));
    const int cw = f2i(floorf(x)), ch = f2i(floorf(y));
    const int h0 = S2C_MAX(ch - Ri, 0), h1 = S2C_MIN(ch + Ri, H - 1) + 1;
    const int w0 = S2C_MAX(cw - Ri, 0), w1 = S2C_MIN(cw + Ri, W - 1) + 1;
    float dc = 0.f;
    for (int h = h0; h < h1; ++h)
        for (int w = w0; w < w1; ++w) {
            const float c0 = bit_of(bs, b, w, h, k, H, W, S);
            const float c1 = bit_of(tg, b, w, h, k, H, W, S);
            const float wt = gauss_w(x, y, w, h, R);
            dc += wt * (c0 - c1) * (c1 == 0.f ? rho : 1.f);
        }
    return dc;
}

/* antialias.cu:179-279 (serial accumulation in fragment order). */
void orc_backward_antialiased_cluster_bitset(
    int N, const int *frag_pix, const float *fa, int H, int W, int S,
    const int *bitset, const int *target, int V, const float *pos, int E,
    const int *edges, const float *normals, const int *edge2cluster,
    float *d_pos, float R, float rho, float *absum) {
    const float fH = (float)H, fW = (float)W;
    for (int i = 0; i < N; ++i) {
        const int b = frag_pix[i * 3 + 0];
        int h = frag_pix[i * 3 + 1], w = frag_pix[i * 3 + 2];
        const float axis = fa[(size_t)i * 4 + 0];
        const float t1 = fa[(size_t)i * 4 + 1];
        const float t0 = 1.f - t1;
        const int e = f2i(fa[(size_t)i * 4 + 3]) - 1;
        if (e < 0) continue;
        const float *p = pos + (size_t)b * V * 4;
        const float *nb = normals + (size_t)b * E * 2;
        float *dp = d_pos + (size_t)b * V * 4;
        const int v0 = edges[e * 2 + 0], v1 = edges[e * 2 + 1];
        const float p0x = p[v0 * 4 + 0], p0y = p[v0 * 4 + 1], p0w = p[v0 * 4 + 3];
        const float p1x = p[v1 * 4 + 0], p1y = p[v1 * 4 + 1], p1w = p[v1 * 4 + 3];
        const float px = t0 * p0x + t1 * p1x;
        const float py = t0 * p0y + t1 * p1y;
        const float pw = t0 * p0w + t1 * p1w;
        const float fx = 0.5f * (1.f + px / pw) * fW;
        const float fy = 0.5f * (1.f - py / pw) * fH;
        const float nx = nb[(size_t)e * 2 + 0], ny = nb[(size_t)e * 2 + 1];
        int nh = h, nw = w;
        const float fxr = fx - floorf(fx);
        const float fyr = fy - floorf(fy);
        const float eps = 1e-4f;
        float dot = 0.f;
        if (axis == 0.f) {
            if (!isfinite(fxr) || fabsf(fxr - 0.5f) >= eps) continue;
            const int d = 2 * (int)(fyr > 0.5f) - 1;
            nh += d;
            dot += (float)d * ny;
        } else {
            if (!isfinite(fyr) || fabsf(fyr - 0.5f) >= eps) continue;
            const int d = 2 * (int)(fxr > 0.5f) - 1;
            nw += d;
            dot += (float)d * nx;
        }
        if (nw < 0 || nw >= W || nh < 0 || nh >= H || dot == 0.f) continue;
        const int k = edge2cluster[e];
        const float c01 = (dot > 0.f) ? 1.f - bit_of(bitset, b, nw, nh, k, H, W, S)
                                      : 1.f - bit_of(bitset, b, w, h, k, H, W, S);
        const float dc = c01 * adjoint_sum(b, fx, fy, k, H, W, S, bitset, target, R, rho);
        if (dc == 0.f) continue;
        const float dx = dc * 0.5f * fW * nx;
        const float dy = dc * -0.5f * fH * ny;
        const float dpx = dx / pw;
        const float dpy = dy / pw;
        const float dpw = (dx * px + dy * py) / (pw * pw);
        if (isfinite(dpx) && isfinite(dpy) && isfinite(dpw)) {
            dp[v0 * 4 + 0] += dpx * t0; dp[v0 * 4 + 1] += dpy * t0; dp[v0 * 4 + 3] += dpw * t0;
            dp[v1 * 4 + 0] += dpx * t1; dp[v1 * 4 + 1] += dpy * t1; dp[v1 * 4 + 3] += dpw * t1;
            if (absum) {
                /* dpw's own numerator cancels: its rounding scales with
                 * |dx px| + |dy py|, not with |dpw| */
                const float mw = (fabsf(dx * px) + fabsf(dy * py)) / (pw * pw);
                float *ab = absum + (size_t)b * V * 4;
                ab[v0 * 4 + 0] += fabsf(dpx * t0); ab[v0 * 4 + 1] += fabsf(dpy * t0);
                ab[v0 * 4 + 3] += mw * t0;
                ab[v1 * 4 + 0] += fabsf(dpx * t1); ab[v1 * 4 + 1] += fabsf(dpy * t1);
                ab[v1 * 4 + 3] += mw * t1;
            }
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Strand smoothing / skinning — curve_utils.cu                              */
/* ------------------------------------------------------------------------ */

/* curve_utils.cu:24-68.  Output entries with total weight <= 1e-6 are left
 * as the caller filled them (zeros, curve_utils.py:203). */
void orc_smooth_strands(int N, int P, const float *x, const float *s,
                        float *out, float beta, float *absum) {
    const float inv_b2 = 1.0f / (beta * beta);
    const float cut = 3.0f * beta;
    for (int n = 0; n < N; ++n) {
        const float *xs = x + (size_t)n * P * 3;
        const float *ss = s + (size_t)n * P;
        for (int i = 0; i < P; ++i) {
            float a0 = 0.f, a1 = 0.f, a2 = 0.f, sw = 0.f;
            double m0 = 0., m1 = 0., m2 = 0.;  /* sums of |w x| for `absum` */
            for (int j = 0; j < P; ++j) {
                const float dist = fabsf(ss[i] - ss[j]);
                if (dist > cut) continue;
                const float w = expf(-(dist * dist) * inv_b2);
                a0 += w * xs[j * 3 + 0]; a1 += w * xs[j * 3 + 1]; a2 += w * xs[j * 3 + 2];
                sw += w;
                m0 += fabs((double)w * xs[j * 3 + 0]);
                m1 += fabs((double)w * xs[j * 3 + 1]);
                m2 += fabs((double)w * xs[j * 3 + 2]);
            }
            if (sw > 1e-6f) {
                float *o = out + ((size_t)n * P + i) * 3;
                o[0] = a0 / sw; o[1] = a1 / sw; o[2] = a2 / sw;
                if (absum) {
                    float *ab = absum + ((size_t)n * P + i) * 3;
                    ab[0] = (float)(m0 / sw); ab[1] = (float)(m1 / sw); ab[2] = (float)(m2 / sw);
                }
            }
        }
    }
}

/* curve_utils.cu:70-131. */
void orc_skinning(int N, int P, const float *x, const float *s,
                  const float *hnd, const float *can, const float *R,
                  float *out, float beta, float *absum) {
    const float inv_b2 = 1.0f / (beta * beta);
    const float cut = 3.0f * beta;
    for (int n = 0; n < N; ++n) {
        const size_t o3 = (size_t)n * P * 3;
        for (int i = 0; i < P; ++i) {
            const float si = s[(size_t)n * P + i];
            float a0 = 0.f, a1 = 0.f, a2 = 0.f, sw = 0.f;
            double m[3] = {0., 0., 0.};  /* weighted sums of |term| for `absum` */
            const float *xi = x + o3 + (size_t)i * 3;
            for (int j = 0; j < P; ++j) {
                const float dist = fabsf(si - s[(size_t)n * P + j]);
                if (dist > cut) continue;
                const float w = expf(-(dist * dist) * inv_b2);
                const float *hj = hnd + o3 + (size_t)j * 3;
                const float *cj = can + o3 + (size_t)j * 3;
                const float *Rj = R + ((size_t)n * P + j) * 9;
                const float l0 = xi[0] - hj[0], l1 = xi[1] - hj[1], l2 = xi[2] - hj[2];
                const float r0 = Rj[0] * l0 + Rj[1] * l1 + Rj[2] * l2;
                const float r1 = Rj[3] * l0 + Rj[4] * l1 + Rj[5] * l2;
                const float r2 = Rj[6] * l0 + Rj[7] * l1 + Rj[8] * l2;
                a0 += w * (r0 + cj[0]); a1 += w * (r1 + cj[1]); a2 += w * (r2 + cj[2]);
                sw += w;
                for (int d = 0; d < 3; ++d)
                    m[d] += (double)w * (fabs((double)Rj[3 * d] * l0) + fabs((double)Rj[3 * d + 1] * l1) +
                                         fabs((double)Rj[3 * d + 2] * l2) + fabs((double)cj[d]) +
                                         fabs((double)l0) + fabs((double)l1) + fabs((double)l2));
            }
            float *o = out + o3 + (size_t)i * 3;
            float *ab = absum ? absum + o3 + (size_t)i * 3 : NULL;
            if (sw > 1e-8f) {
                o[0] = a0 / sw; o[1] = a1 / sw; o[2] = a2 / sw;
                if (ab) { ab[0] = (float)(m[0] / sw); ab[1] = (float)(m[1] / sw); ab[2] = (float)(m[2] / sw); }
            } else {
                o[0] = xi[0]; o[1] = xi[1]; o[2] = xi[2];
                if (ab) ab[0] = ab[1] = ab[2] = 0.f;
            }
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Composite forms (the reference's call sequences, as one CPU function)     */
/* ------------------------------------------------------------------------ */

/* dda.py:5-76 followed by rasterize.py:170-214 (3_fitting.py:254-261):
 * every edge of every view rasterized by DDA and OR-ed into a zero-filled
 * (B, H, W, S) bitset.  `threads` > 1 splits views x edge ranges over
 * OpenMP threads (OR is order-free, so the result is the same). */
void orc_dda_cluster_mask(int B, int H, int W, int V, const float *pos, int E,
                          const int *edges, const int *edge2cluster, int S,
                          int *bitset, int threads) {
    const long long P = (long long)B * E;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 4096) num_threads(threads > 0 ? threads : 1)
#endif
    for (long long q = 0; q < P; ++q) {
        float sl[2], sp[4];
        const int id = (int)q;
        const int cnt = dda_span_one(id, H, W, V, pos, E, edges, sl, sp);
        const int b = id / E, e = id % E;
        for (int k = 0; k < cnt; ++k) {
            int pix[3];
            float at[4];
            if (!dda_frag_one(k, b, e, sl, sp, H, W, V, pos, edges, pix, at)) continue;
            const int c = edge2cluster[e];
            const long long idx = (((long long)b * H + pix[1]) * W + pix[2]) * S + c / 32;
            __atomic_fetch_or((uint32_t *)&bitset[idx], 1u << (c % 32), __ATOMIC_RELAXED);
        }
    }
    (void)threads;
}

/* rasterize.py:6-53 followed by rasterize.py:170-214 (3_fitting.py:154-163):
 * triangle fragments OR-ed into a zero-filled bitset. */
void orc_triangle_cluster_mask(int B, int H, int W, int V, const float *pos,
                               int T, const int *tri, const int *tri2cluster,
                               int S, int *bitset) {
    for (int i = 0; i < B * T; ++i) {
        int r[4];
        const int b = i / T, t = i % T;
        const int cnt = tri_rect_one(b, t, H, W, V, pos, tri, r);
        for (int l = 0; l < cnt; ++l) {
            int pix[3];
            float at[4];
            if (!tri_frag_one(b, t, r[2] + l % r[3], r[0] + l / r[3], H, W, V, pos,
                              tri, pix, at))
                continue;
            const int c = tri2cluster[t];
            const long long idx = (((long long)b * H + pix[1]) * W + pix[2]) * S + c / 32;
            bitset[idx] |= (int)(1u << (c % 32));
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Area-sampled card coverage (the differentiable card-fitting form)         */
/* ------------------------------------------------------------------------ */
/* The reference differentiates its rasterize + antialias chain
 * (3_fitting.py:380-400) through a surrogate edge gradient.  The adjoint
 * form of this library renders area-sampled coverage instead,
 *   C[b, h, w] = sum_t area(triangle t of view b ∩ [w, w+1] x [h, h+1]),
 * in the screen coordinates of dda.cu:33-36, so the gradient is exact and
 * checkable by finite differences.  This restatement works in double and
 * uses a different algorithm from the HIP kernel: the pixel square is
 * clipped by the triangle's three half-planes (Sutherland-Hodgman) and the
 * polygon's area is taken with the shoelace formula.  The backward is the
 * transport theorem: moving the vertices of edge a->b changes the covered
 * area by the outward normal times |ab| integrated against (1 - t) (for a)
 * and t (for b) over the part of the edge inside the pixel box. */

typedef struct {
    double p[3][2];
    int map[3];
    int valid;
} CovTri;

static CovTri cov_tri(const double *pos, long long vb, const int *t, int H, int W) {
    CovTri c;
    c.valid = 1;
    for (int i = 0; i < 3; ++i) {
        const double *P = pos + (vb + t[i]) * 4;
        if (!(P[3] > 0.0)) c.valid = 0;
        c.p[i][0] = 0.5 * (1.0 + P[0] / P[3]) * W;
        c.p[i][1] = 0.5 * (1.0 - P[1] / P[3]) * H;
        c.map[i] = i;
    }
    const double cr = (c.p[1][0] - c.p[0][0]) * (c.p[2][1] - c.p[0][1]) -
                      (c.p[1][1] - c.p[0][1]) * (c.p[2][0] - c.p[0][0]);
    if (!(cr != 0.0) || !isfinite(cr)) c.valid = 0;
    if (cr < 0.0) {
        double q0 = c.p[1][0], q1 = c.p[1][1];
        c.p[1][0] = c.p[2][0];
        c.p[1][1] = c.p[2][1];
        c.p[2][0] = q0;
        c.p[2][1] = q1;
        c.map[1] = 2;
        c.map[2] = 1;
    }
    return c;
}

/* Pixel-rectangle of the triangle, clamped to the image: r = {h0, rows, w0, cols}. */
static void cov_rect(const CovTri *c, int H, int W, int r[4]) {
    double xmin = c->p[0][0], xmax = xmin, ymin = c->p[0][1], ymax = ymin;
    for (int i = 1; i < 3; ++i) {
        xmin = fmin(xmin, c->p[i][0]);
        xmax = fmax(xmax, c->p[i][0]);
        ymin = fmin(ymin, c->p[i][1]);
        ymax = fmax(ymax, c->p[i][1]);
    }
    const double w0 = fmax(floor(xmin), 0.0), w1 = fmin(floor(xmax), W - 1.0);
    const double h0 = fmax(floor(ymin), 0.0), h1 = fmin(floor(ymax), H - 1.0);
    r[0] = (int)h0;
    r[1] = h1 >= h0 ? (int)(h1 - h0) + 1 : 0;
    r[2] = (int)w0;
    r[3] = w1 >= w0 ? (int)(w1 - w0) + 1 : 0;
    if (!c->valid) r[1] = r[3] = 0;
}

/* Area of the unit pixel square at (bx, by) inside the positively oriented
 * triangle. */
static double square_in_tri(const CovTri *c, double bx, double by) {
    double a[16][2], b[16][2];
    int n = 4;
    const double sq[4][2] = {{0, 0}, {1, 0}, {1, 1}, {0, 1}};
    for (int i = 0; i < 4; ++i) {
        a[i][0] = sq[i][0];
        a[i][1] = sq[i][1];
    }
    for (int e = 0; e < 3 && n > 0; ++e) {
        const double ox = c->p[e][0] - bx, oy = c->p[e][1] - by;
        const double dx = c->p[(e + 1) % 3][0] - c->p[e][0];
        const double dy = c->p[(e + 1) % 3][1] - c->p[e][1];
        int m = 0;
        for (int i = 0; i < n; ++i) {
            const double *u = a[i], *v = a[(i + 1) % n];
            const double fu = dx * (u[1] - oy) - dy * (u[0] - ox);
            const double fv = dx * (v[1] - oy) - dy * (v[0] - ox);
            if (fu >= 0) {
                b[m][0] = u[0];
                b[m][1] = u[1];
                ++m;
            }
            if ((fu >= 0) != (fv >= 0)) {
                const double s = fu / (fu - fv);
                b[m][0] = u[0] + s * (v[0] - u[0]);
                b[m][1] = u[1] + s * (v[1] - u[1]);
                ++m;
            }
        }
        n = m;
        for (int i = 0; i < n; ++i) {
            a[i][0] = b[i][0];
            a[i][1] = b[i][1];
        }
    }
    double s = 0;
    for (int i = 0; i < n; ++i) s += a[i][0] * a[(i + 1) % n][1] - a[i][1] * a[(i + 1) % n][0];
    return 0.5 * s;
}

/* out[(b*h + i)*w + j] += coverage of pixel (y0[b] + i, x0[b] + j) by the
 * triangles sel[0..n_sel) (all T triangles when sel is NULL) of view b. */
void orc_card_coverage(int B, int H, int W, int V, const double *pos, int T, const int *tri,
                       int n_sel, const int *sel, const int *y0, const int *x0, int h, int w,
                       double *out) {
    const int n = sel ? n_sel : T;
    for (int b = 0; b < B; ++b) {
        const int oy = y0 ? y0[b] : 0, ox = x0 ? x0[b] : 0;
        for (int k = 0; k < n; ++k) {
            const int t = sel ? sel[k] : k;
            const CovTri c = cov_tri(pos, (long long)b * V, tri + 3 * t, H, W);
            int r[4];
            cov_rect(&c, H, W, r);
            for (int y = r[0]; y < r[0] + r[1]; ++y) {
                if (y < oy || y >= oy + h) continue;
                for (int x = r[2]; x < r[2] + r[3]; ++x) {
                    if (x < ox || x >= ox + w) continue;
                    out[((long long)b * h + (y - oy)) * w + (x - ox)] += square_in_tri(&c, x, y);
                }
            }
        }
    }
}

/* d_pos[b, v, (x, y, w)] += d loss / d pos given d_cov = d loss / d C. */
void orc_card_coverage_grad(int B, int H, int W, int V, const double *pos, int T,
                            const int *tri, const double *d_cov, double *d_pos) {
    for (int b = 0; b < B; ++b) {
        for (int t = 0; t < T; ++t) {
            const CovTri c = cov_tri(pos, (long long)b * V, tri + 3 * t, H, W);
            int r[4];
            cov_rect(&c, H, W, r);
            double gs[3][2] = {{0, 0}, {0, 0}, {0, 0}};
            for (int y = r[0]; y < r[0] + r[1]; ++y) {
                for (int x = r[2]; x < r[2] + r[3]; ++x) {
                    const double g = d_cov[((long long)b * H + y) * W + x];
                    if (g == 0.0) continue;
                    for (int e = 0; e < 3; ++e) {
                        const int f = (e + 1) % 3;
                        const double px = c.p[e][0], py = c.p[e][1];
                        const double dx = c.p[f][0] - px, dy = c.p[f][1] - py;
                        /* Liang-Barsky clip of p + s d to the box */
                        const double pp[4] = {-dx, dx, -dy, dy};
                        const double qq[4] = {px - x, x + 1.0 - px, py - y, y + 1.0 - py};
                        double t0 = 0.0, t1 = 1.0;
                        int empty = 0;
                        for (int i = 0; i < 4; ++i) {
                            if (pp[i] == 0.0) {
                                if (qq[i] < 0.0) empty = 1;
                            } else if (pp[i] < 0.0) {
                                t0 = fmax(t0, qq[i] / pp[i]);
                            } else {
                                t1 = fmin(t1, qq[i] / pp[i]);
                            }
                        }
                        if (empty || !(t1 > t0)) continue;
                        const double i1 = 0.5 * (t1 * t1 - t0 * t0), i0 = (t1 - t0) - i1;
                        gs[e][0] += g * dy * i0;
                        gs[e][1] += -g * dx * i0;
                        gs[f][0] += g * dy * i1;
                        gs[f][1] += -g * dx * i1;
                    }
                }
            }
            for (int i = 0; i < 3; ++i) {
                const long long row = (long long)b * V + tri[3 * t + c.map[i]];
                const double *P = pos + row * 4;
                const double sx = 0.5 * W, sy = 0.5 * H;
                d_pos[row * 4 + 0] += gs[i][0] * sx / P[3];
                d_pos[row * 4 + 1] += -gs[i][1] * sy / P[3];
                d_pos[row * 4 + 3] += (-gs[i][0] * sx * P[0] + gs[i][1] * sy * P[1]) / (P[3] * P[3]);
            }
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Bishop frames (curve_utils.py:76-165, torch ops restated in fp32)        */
/* ------------------------------------------------------------------------ */
/* Tangent of vertex i = normalized (x[i+1]-x[i]) + (x[i]-x[i-1]) (one-sided at
 * the ends, curve_utils.py:120-125); root normal/binormal from the
 * branchless orthonormal basis of tangent_frame (curve_utils.py:76-102);
 * then the double-reflection recurrence (curve_utils.py:140-163).  Norms
 * are clamped at 1e-8 and squared lengths at 1e-8, as there. */
static void v3_unit(float *v) {
    float n = sqrtf(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    if (n < 1e-8f) n = 1e-8f;
    v[0] /= n;
    v[1] /= n;
    v[2] /= n;
}

void orc_bishop_frame(int N, int S, const float *x, float *T, float *Nr, float *Bn) {
    for (int n = 0; n < N; ++n) {
        const float *p = x + (long long)n * S * 3;
        float *t = T + (long long)n * S * 3, *nr = Nr + (long long)n * S * 3,
              *bn = Bn + (long long)n * S * 3;
        for (int i = 0; i < S; ++i) {
            float v[3] = {0, 0, 0};
            for (int k = 0; k < 3; ++k) {
                if (i + 1 < S) v[k] += p[3 * (i + 1) + k] - p[3 * i + k];
                if (i > 0) v[k] += p[3 * i + k] - p[3 * (i - 1) + k];
            }
            v3_unit(v);
            memcpy(t + 3 * i, v, sizeof v);
        }
        if (S == 0) continue;
        float t0[3] = {t[0], t[1], t[2]};
        v3_unit(t0);
        const float sg = t0[2] < 0.f ? -1.f : 1.f;
        const float a = -1.f / (sg + t0[2]);
        const float b = t0[0] * t0[1] * a;
        nr[0] = 1.f + sg * t0[0] * t0[0] * a;
        nr[1] = sg * b;
        nr[2] = -sg * t0[0];
        bn[0] = b;
        bn[1] = sg + t0[1] * t0[1] * a;
        bn[2] = -t0[1];
        for (int i = 1; i < S; ++i) {
            float v1[3], nL[3], tL[3], v2[3], n1[3], b1[3];
            for (int k = 0; k < 3; ++k) v1[k] = p[3 * i + k] - p[3 * (i - 1) + k];
            float c1 = v1[0] * v1[0] + v1[1] * v1[1] + v1[2] * v1[2];
            if (c1 < 1e-8f) c1 = 1e-8f;
            const float *np_ = nr + 3 * (i - 1), *tp = t + 3 * (i - 1), *tc = t + 3 * i;
            const float dn = v1[0] * np_[0] + v1[1] * np_[1] + v1[2] * np_[2];
            const float dt = v1[0] * tp[0] + v1[1] * tp[1] + v1[2] * tp[2];
            for (int k = 0; k < 3; ++k) {
                nL[k] = np_[k] - 2.f * (dn / c1) * v1[k];
                tL[k] = tp[k] - 2.f * (dt / c1) * v1[k];
                v2[k] = tc[k] - tL[k];
            }
            float c2 = v2[0] * v2[0] + v2[1] * v2[1] + v2[2] * v2[2];
            if (c2 < 1e-8f) c2 = 1e-8f;
            const float d2 = v2[0] * nL[0] + v2[1] * nL[1] + v2[2] * nL[2];
            for (int k = 0; k < 3; ++k) n1[k] = nL[k] - 2.f * (d2 / c2) * v2[k];
            v3_unit(n1);
            b1[0] = tc[1] * n1[2] - tc[2] * n1[1];
            b1[1] = tc[2] * n1[0] - tc[0] * n1[2];
            b1[2] = tc[0] * n1[1] - tc[1] * n1[0];
            v3_unit(b1);
            memcpy(nr + 3 * i, n1, sizeof n1);
            memcpy(bn + 3 * i, b1, sizeof b1);
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Arclength resampling (hair_loader.cpp:261-307)                            */
/* ------------------------------------------------------------------------ */
/* Per strand: cumulative float lengths divided by the total; sample j < T-1
 * at t = j/(T-1) interpolates points idx-1, idx around the upper bound of t
 * (clamped to [1, N-1]); the last sample copies the last point. */
static int upper_bound_f(const float *l, int n, float t) {
    int first = 0, len = n;
    while (len > 0) {
        int half = len >> 1;
        if (t < l[first + half]) {
            len = half;
        } else {
            first += half + 1;
            len -= half + 1;
        }
    }
    return first;
}

void orc_resample_strands_by_arclength(int num_strands, const int *num_samples, const float *points,
                                       int T, float *out) {
    long long off = 0;
    for (int i = 0; i < num_strands; ++i) {
        const int N = num_samples[i];
        const float *p = points + off * 3;
        float *l = (float *)malloc(sizeof(float) * (size_t)N);
        l[0] = 0.f;
        for (int j = 1; j < N; ++j) {
            const float dx = p[j * 3 + 0] - p[(j - 1) * 3 + 0];
            const float dy = p[j * 3 + 1] - p[(j - 1) * 3 + 1];
            const float dz = p[j * 3 + 2] - p[(j - 1) * 3 + 2];
            l[j] = l[j - 1] + sqrtf(dx * dx + dy * dy + dz * dz);
        }
        const float total = l[N - 1];
        for (int j = 0; j < N; ++j) l[j] /= total;
        float *o = out + (long long)i * T * 3;
        for (int j = 0; j < T - 1; ++j) {
            const float t = (float)j / (float)(T - 1);
            int idx = upper_bound_f(l, N, t);
            idx = S2C_MIN(S2C_MAX(idx, 1), N - 1);
            const float t0 = (l[idx] - t) / (l[idx] - l[idx - 1]);
            const float t1 = (t - l[idx - 1]) / (l[idx] - l[idx - 1]);
            for (int k = 0; k < 3; ++k) o[j * 3 + k] = t0 * p[(idx - 1) * 3 + k] + t1 * p[idx * 3 + k];
        }
        for (int k = 0; k < 3; ++k) o[(T - 1) * 3 + k] = p[(N - 1) * 3 + k];
        free(l);
        off += N;
    }
}
