// Card mesh inputs built on the GPU, so that fitting jobs need no host
// preparation:
//   * strand -> strip sweep (2_texture.py:17-96): double-reflection
//     rotation-minimizing frames, one lane per (strand, rotation);
//   * grid-mesh topology (texture_utils.py:525-655): faces, quads, uvs,
//     h_edges, w_edges and face clusters of concatenated wisps;
//   * strand polyline edges with their cluster (3_fitting.py:17-60).
// Float arithmetic is fp32 in the reference's order (its arrays are
// float32); the root-frame rotation, which the reference evaluates with a
// float64 angle, is evaluated in double and rounded, as the reference
// rounds it into its float32 frame.
#include "common.hpp"

namespace s2c {

constexpr int kMaxStripRotations = 16;

struct StripRotations {
    double c[kMaxStripRotations], s[kMaxStripRotations];
};

__device__ __forceinline__ float3 f3(const float* p) { return make_float3(p[0], p[1], p[2]); }
__device__ __forceinline__ float3 sub3(float3 a, float3 b) {
    return make_float3(a.x - b.x, a.y - b.y, a.z - b.z);
}
__device__ __forceinline__ float3 add3(float3 a, float3 b) {
    return make_float3(a.x + b.x, a.y + b.y, a.z + b.z);
}
__device__ __forceinline__ float3 mul3(float k, float3 a) {
    return make_float3(k * a.x, k * a.y, k * a.z);
}
__device__ __forceinline__ float3 div3(float3 a, float k) {
    return make_float3(a.x / k, a.y / k, a.z / k);
}
__device__ __forceinline__ float dot3(float3 a, float3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
__device__ __forceinline__ float norm3(float3 a) { return sqrtf(dot3(a, a)); }
__device__ __forceinline__ float3 cross3(float3 a, float3 b) {
    return make_float3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}

// Central-difference tangent at point i of a strand of n >= 2 points.
__device__ __forceinline__ float3 strip_tangent(const float* p, int i, int n) {
    float3 t = make_float3(0.f, 0.f, 0.f);
    if (i < n - 1) t = add3(t, sub3(f3(p + 3 * (i + 1)), f3(p + 3 * i)));
    if (i > 0) t = add3(t, sub3(f3(p + 3 * i), f3(p + 3 * (i - 1))));
    return div3(t, fmaxf(norm3(t), 1e-8f));
}

// One lane per (strand, rotation): the strip of res_w rows written at
// out + 3 * res_w * (num_rot * offsets[s] + r * counts[s]), row i at
// offset -width/2 + i * width/(res_w - 1) along the binormal.  A strand of
// fewer than 2 points has no tangent and is left unwritten (the reference
// indexes its second tangent).
__global__ void __launch_bounds__(kBlock)
strand_strip_kernel(int num_strands, const int* __restrict__ counts,
                    const long long* __restrict__ offsets, const float* __restrict__ points,
                    int res_w, double width, int num_rot, StripRotations rot,
                    float* __restrict__ out) {
    const long long tidx = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x;
    if (tidx >= static_cast<long long>(num_strands) * num_rot) return;
    const int s = static_cast<int>(tidx / num_rot), r = static_cast<int>(tidx % num_rot);
    const int n = counts[s];
    if (n < 2) return;
    const float* p = points + 3 * offsets[s];
    float* o = out + 3ll * res_w * (num_rot * offsets[s] + static_cast<long long>(r) * n);
    // root frame (2_texture.py:35-51)
    const float3 T0 = strip_tangent(p, 0, n);
    const float3 dT = sub3(strip_tangent(p, 1, n), T0);
    float3 nb;
    if (norm3(dT) < 1e-4f) {
        nb = make_float3(0.f, 1.f, 0.f);
        if (fabsf(dot3(T0, nb)) > 0.9f) nb = make_float3(0.f, 0.f, 1.f);
    } else {
        nb = mul3(-1.f, div3(dT, norm3(dT)));
    }
    float3 bb = cross3(T0, nb);
    bb = div3(bb, fmaxf(norm3(bb), 1e-8f));
    nb = cross3(bb, T0);
    // rotation about T0 (Rodrigues, 2_texture.py:61-67), in double
    const double c = rot.c[r], sn = rot.s[r];
    const double tn = static_cast<double>(dot3(T0, nb));
    const float3 tx = cross3(T0, nb);
    const float3 N0 = make_float3(
        static_cast<float>(c * nb.x + sn * tx.x + (1.0 - c) * tn * T0.x),
        static_cast<float>(c * nb.y + sn * tx.y + (1.0 - c) * tn * T0.y),
        static_cast<float>(c * nb.z + sn * tx.z + (1.0 - c) * tn * T0.z));
    // the reference's B0 = cross(T0, N0) with N0 still in double
    const double n0x = c * nb.x + sn * tx.x + (1.0 - c) * tn * T0.x;
    const double n0y = c * nb.y + sn * tx.y + (1.0 - c) * tn * T0.y;
    const double n0z = c * nb.z + sn * tx.z + (1.0 - c) * tn * T0.z;
    float3 B = make_float3(static_cast<float>(T0.y * n0z - T0.z * n0y),
                           static_cast<float>(T0.z * n0x - T0.x * n0z),
                           static_cast<float>(T0.x * n0y - T0.y * n0x));
    float3 N = N0;
    // row offsets: numpy's linspace(-width/2, width/2, res_w) in double,
    // rounded to float32 (2_texture.py:92)
    const double lo = -0.5 * width, hi = 0.5 * width;
    const double step = (hi - lo) / static_cast<double>(res_w - 1);
    auto emit = [&](int i, float3 b) {
        const float3 q = f3(p + 3 * i);
        for (int w = 0; w < res_w; ++w) {
            const float off = static_cast<float>(w == res_w - 1 ? hi : lo + static_cast<double>(w) * step);
            float* d = o + 3ll * (static_cast<long long>(w) * n + i);
            d[0] = q.x + off * b.x;
            d[1] = q.y + off * b.y;
            d[2] = q.z + off * b.z;
        }
    };
    emit(0, B);
    float3 tprev = T0;
    for (int i = 1; i < n; ++i) {
        const float3 ti = strip_tangent(p, i, n);
        float3 v = sub3(f3(p + 3 * i), f3(p + 3 * (i - 1)));
        const float vn = norm3(v);
        if (!(vn < 1e-8f)) {  // double reflection (2_texture.py:73-87)
            v = div3(v, vn);
            const float3 nl = sub3(N, mul3(2.f * dot3(N, v), v));
            const float3 tl = sub3(tprev, mul3(2.f * dot3(tprev, v), v));
            float3 v2 = sub3(ti, tl);
            v2 = div3(v2, fmaxf(norm3(v2), 1e-8f));
            float3 n1 = sub3(nl, mul3(2.f * dot3(nl, v2), v2));
            n1 = div3(n1, fmaxf(norm3(n1), 1e-8f));
            N = n1;
            B = cross3(ti, n1);
        }
        emit(i, B);
        tprev = ti;
    }
}

// One workgroup per grid: faces (2 per quad, [a, b, c], [a, c, d] of quad
// [a, b, c, d] = [(i,j), (i+1,j), (i+1,j+1), (i,j+1)]), quads, h_edges,
// w_edges, face clusters and uvs (u over rows, v over columns, numpy's
// linspace in double: start + k * (stop - start) / (num - 1), the last
// sample exactly stop).
__global__ void __launch_bounds__(kBlock)
grid_topology_kernel(const int* __restrict__ rows, const int* __restrict__ cols,
                     const int* __restrict__ voff, const int* __restrict__ qoff,
                     const int* __restrict__ hoff, const int* __restrict__ woff,
                     const long long* __restrict__ grid2cluster, const double* __restrict__ rects,
                     double margin_top, int* __restrict__ faces, int* __restrict__ quads,
                     int* __restrict__ h_edges, int* __restrict__ w_edges,
                     long long* __restrict__ face2cluster, double* __restrict__ uvs) {
    const int g = blockIdx.x;
    const int S = rows[g], N = cols[g], v0 = voff[g];
    const int nq = (S - 1) * (N - 1);
    for (int q = threadIdx.x; q < nq; q += kBlock) {
        const int i = q / (N - 1), j = q - i * (N - 1);
        const int a = v0 + i * N + j, b = a + N, c = b + 1, d = a + 1;
        int* qq = quads + 4ll * (qoff[g] + q);
        qq[0] = a; qq[1] = b; qq[2] = c; qq[3] = d;
        int* f = faces + 6ll * (qoff[g] + q);
        f[0] = a; f[1] = b; f[2] = c;
        f[3] = a; f[4] = c; f[5] = d;
        if (face2cluster) {
            face2cluster[2ll * (qoff[g] + q)] = grid2cluster[g];
            face2cluster[2ll * (qoff[g] + q) + 1] = grid2cluster[g];
        }
    }
    const int nh = S * (N - 1);
    for (int e = threadIdx.x; e < nh; e += kBlock) {
        const int i = e / (N - 1), j = e - i * (N - 1);
        h_edges[2ll * (hoff[g] + e)] = v0 + i * N + j;
        h_edges[2ll * (hoff[g] + e) + 1] = v0 + i * N + j + 1;
    }
    const int nw = (S - 1) * N;
    for (int e = threadIdx.x; e < nw; e += kBlock) {
        w_edges[2ll * (woff[g] + e)] = v0 + e;
        w_edges[2ll * (woff[g] + e) + 1] = v0 + e + N;
    }
    if (uvs) {
        const double u0 = rects[4 * g], vv0 = rects[4 * g + 1], u1 = rects[4 * g + 2];
        const double v1 = rects[4 * g + 3] - margin_top;
        auto lin = [](double a, double b, int num, int k) {
            if (num == 1) return a;
            if (k == num - 1) return b;
            const double step = (b - a) / static_cast<double>(num - 1);
            return a + static_cast<double>(k) * step;
        };
        for (int v = threadIdx.x; v < S * N; v += kBlock) {
            const int i = v / N, j = v - i * N;
            uvs[2ll * (v0 + v)] = lin(u0, u1, S, i);
            uvs[2ll * (v0 + v) + 1] = lin(vv0, v1, N, j);
        }
    }
}

// Edges between consecutive points of each strand, with its cluster.
__global__ void __launch_bounds__(kBlock)
strand_edges_kernel(int num_strands, const int* __restrict__ counts,
                    const long long* __restrict__ offsets, const int* __restrict__ cluster_ids,
                    int* __restrict__ edges, int* __restrict__ edge2cluster) {
    const int s = blockIdx.x;
    if (s >= num_strands) return;
    const int n = counts[s];
    const long long p0 = offsets[s], e0 = p0 - s;  // strand s's first edge: points before it minus s
    for (int i = threadIdx.x; i < n - 1; i += kBlock) {
        edges[2 * (e0 + i)] = static_cast<int>(p0 + i);
        edges[2 * (e0 + i) + 1] = static_cast<int>(p0 + i + 1);
        edge2cluster[e0 + i] = cluster_ids[s];
    }
}

}  // namespace s2c

using namespace s2c;

extern "C" int s2c_strands_to_strips(int num_strands, const int* counts, const long long* offsets,
                                     const float* points, int res_w, double width, int num_rot,
                                     const float* rot_deg, float* out, void* stream) {
    S2C_CHECK(num_strands >= 0, "strands_to_strips: num_strands must be >= 0");
    S2C_CHECK(res_w >= 2, "strands_to_strips: res_w must be at least 2");
    S2C_CHECK(num_rot >= 1 && num_rot <= kMaxStripRotations,
              "strands_to_strips: 1 to %d tangent rotations", kMaxStripRotations);
    S2C_CHECK(rot_deg != nullptr, "strands_to_strips: null rotations");
    if (num_strands == 0) return S2C_OK;
    StripRotations rot{};
    for (int r = 0; r < num_rot; ++r) {
        const double th = static_cast<double>(rot_deg[r]) * (3.14159265358979323846 / 180.0);
        rot.c[r] = cos(th);
        rot.s[r] = sin(th);
    }
    hipStream_t s = static_cast<hipStream_t>(stream);
    const long long lanes = static_cast<long long>(num_strands) * num_rot;
    strand_strip_kernel<<<grid_for(lanes, kBlock), kBlock, 0, s>>>(
        num_strands, counts, offsets, points, res_w, width, num_rot, rot, out);
    S2C_LAUNCH_CHECK();
    return S2C_OK;
}

extern "C" int s2c_grid_mesh_topology(int num_grids, const int* rows, const int* cols,
                                      const int* voff, const int* qoff, const int* hoff,
                                      const int* woff, const long long* grid2cluster,
                                      const double* rects, double margin_top, int* faces,
                                      int* quads, int* h_edges, int* w_edges,
                                      long long* face2cluster, double* uvs, void* stream) {
    S2C_CHECK(num_grids >= 0, "grid_mesh_topology: num_grids must be >= 0");
    S2C_CHECK((face2cluster == nullptr) == (grid2cluster == nullptr),
              "grid_mesh_topology: face2cluster and grid2cluster go together");
    S2C_CHECK((uvs == nullptr) == (rects == nullptr), "grid_mesh_topology: uvs and rects go together");
    if (num_grids == 0) return S2C_OK;
    hipStream_t s = static_cast<hipStream_t>(stream);
    grid_topology_kernel<<<num_grids, kBlock, 0, s>>>(rows, cols, voff, qoff, hoff, woff,
                                                      grid2cluster, rects, margin_top, faces,
                                                      quads, h_edges, w_edges, face2cluster, uvs);
    S2C_LAUNCH_CHECK();
    return S2C_OK;
}

extern "C" int s2c_strand_edges(int num_strands, const int* counts, const long long* offsets,
                                const int* cluster_ids, int* edges, int* edge2cluster,
                                void* stream) {
    S2C_CHECK(num_strands >= 0, "strand_edges: num_strands must be >= 0");
    if (num_strands == 0) return S2C_OK;
    hipStream_t s = static_cast<hipStream_t>(stream);
    strand_edges_kernel<<<num_strands, kBlock, 0, s>>>(num_strands, counts, offsets, cluster_ids,
                                                       edges, edge2cluster);
    S2C_LAUNCH_CHECK();
    return S2C_OK;
}
