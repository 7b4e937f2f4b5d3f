// `_core`: the PyTorch-ROCm binding of the C ABI in include/strns2cards_hip.h.
//
// Same module name, function names and argument order as the reference's
// nanobind module (pystrns2cards/src/pystrns2cards.cpp:204-924), so the
// Python wrappers read like the reference's.  Each function validates
// dtype / rank / contiguity (TypeError, as nanobind's casters do) and
// device / shape agreement (RuntimeError, as the reference's explicit
// checks do), then calls the C ABI on the current HIP stream of the
// tensor's device.  HIP failures raise RuntimeError instead of exit().
#include <torch/extension.h>
#include <ATen/hip/impl/HIPGuardImplMasqueradingAsCUDA.h>
#include <ATen/hip/impl/HIPStreamMasqueradingAsCUDA.h>

#include <pybind11/numpy.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <optional>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/strns2cards_hip.h"

namespace py = pybind11;
using at::Tensor;
// On PyTorch-ROCm, GPU tensors carry the CUDA device type; this guard
// accepts that and switches the current HIP device.
using DeviceGuard = c10::hip::HIPGuardMasqueradingAsCUDA;

namespace {

const char* dtype_name(at::ScalarType t) {
    return t == at::kFloat ? "float32" : t == at::kInt ? "int32" : t == at::kLong ? "int64"
           : t == at::kDouble ? "float64" : t == at::kByte ? "uint8" : "other";
}

// Shape spec: -1 = any extent.
void need(const Tensor& t, const char* fn, const char* name, at::ScalarType dt,
          std::initializer_list<int64_t> shape) {
    if (t.scalar_type() != dt)
        throw py::type_error(std::string(fn) + ": `" + name + "` must be " + dtype_name(dt));
    if (t.dim() != static_cast<int64_t>(shape.size()))
        throw py::type_error(std::string(fn) + ": `" + name + "` must have " +
                             std::to_string(shape.size()) + " dimensions");
    int i = 0;
    for (int64_t s : shape) {
        if (s >= 0 && t.size(i) != s)
            throw py::type_error(std::string(fn) + ": `" + name + "` has a wrong extent in dim " +
                                 std::to_string(i));
        ++i;
    }
    if (!t.is_contiguous())
        throw py::type_error(std::string(fn) + ": `" + name + "` must be C-contiguous");
    if (!t.is_cuda())
        throw std::runtime_error(std::string(fn) + ": all tensors must be GPU tensors");
}

void same_device(const char* fn, std::initializer_list<const Tensor*> ts) {
    const Tensor* first = *ts.begin();
    for (const Tensor* t : ts)
        if (t->device() != first->device())
            throw std::runtime_error(std::string(fn) + ": tensors are on different devices");
}

void check(int rc, const char* fn) {
    if (rc != S2C_OK) throw std::runtime_error(std::string(fn) + ": " + s2c_last_error());
}

void* stream_of(const Tensor& t) {
    return static_cast<void*>(
        c10::hip::getCurrentHIPStreamMasqueradingAsCUDA(t.device().index()).stream());
}

int I(int64_t v, const char* fn) {
    if (v > INT32_MAX) throw std::runtime_error(std::string(fn) + ": size exceeds INT32_MAX");
    return static_cast<int>(v);
}

template <class T> T* P(const Tensor& t) { return static_cast<T*>(t.data_ptr()); }

// ---- host-side strand files (hair_loader.cpp:15-256) ------------------------

using Strand = std::vector<float>;  // x, y, z per point

// Little-endian reader over a whole file; a read past the end returns false
// and leaves the value untouched (the reference's fread on a short file).
struct Bytes {
    std::vector<unsigned char> b;
    size_t at = 0;
    template <class T> bool get(T& v) {
        if (at + sizeof(T) > b.size()) return false;
        std::memcpy(&v, b.data() + at, sizeof(T));
        at += sizeof(T);
        return true;
    }
};

bool read_file(const std::string& path, Bytes& out) {
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) {
        std::fprintf(stderr, "Couldn't open %s\n", path.c_str());
        return false;
    }
    unsigned char buf[1 << 16];
    size_t n;
    while ((n = std::fread(buf, 1, sizeof(buf), f)) > 0) out.b.insert(out.b.end(), buf, buf + n);
    std::fclose(f);
    return true;
}

// hair_loader.cpp:15-58: count, then per strand a count and 7 floats per
// point (position, normal, label); a strand with a segment shorter than
// 1e-4 is dropped.
std::vector<Strand> strands_bin(Bytes& r) {
    std::vector<Strand> out;
    int n = 0;
    r.get(n);
    for (int i = 0; i < n; ++i) {
        int m = 0;
        if (!r.get(m) || m < 0) break;
        Strand s(static_cast<size_t>(m) * 3, 0.f);
        for (int j = 0; j < m; ++j) {
            float rec[7] = {0, 0, 0, 0, 0, 0, 0};
            for (float& v : rec) r.get(v);
            s[3 * j] = rec[0];
            s[3 * j + 1] = rec[1];
            s[3 * j + 2] = rec[2];
        }
        bool valid = true;
        for (int j = 0; j + 1 < m && valid; ++j) {
            const float dx = s[3 * j] - s[3 * j + 3], dy = s[3 * j + 1] - s[3 * j + 4],
                        dz = s[3 * j + 2] - s[3 * j + 5];
            valid = std::sqrt(dx * dx + dy * dy + dz * dz) >= 1e-4f;
        }
        if (valid) out.push_back(std::move(s));
    }
    return out;
}

// hair_loader.cpp:61-93: count, then per strand a count and its points; a
// one-point strand is skipped.
std::vector<Strand> strands_usc(Bytes& r) {
    std::vector<Strand> out;
    int n = 0;
    r.get(n);
    for (int i = 0; i < n; ++i) {
        int m = 0;
        if (!r.get(m) || m < 0) break;
        Strand s(static_cast<size_t>(m) * 3, 0.f);
        for (float& v : s) r.get(v);
        if (m != 1) out.push_back(std::move(s));
    }
    return out;
}

// hair_loader.cpp:96-217: the "HAIR" header layout (segments stored as
// count - 1, or one default count; the header is skipped as the reference
// skips it) or the headerless ZJU layout (counts, then points).
std::vector<Strand> strands_hair(Bytes& r) {
    std::vector<Strand> out;
    if (r.b.size() < 4) {
        std::fprintf(stderr, "Failed to read header.\n");
        return out;
    }
    const bool hair = std::memcmp(r.b.data(), "HAIR", 4) == 0;
    unsigned n = 0, count = 0;
    std::vector<unsigned> seg;
    if (hair) {
        r.at = 4;
        unsigned bits = 0, def = 0;
        r.get(n);
        r.get(count);
        r.get(bits);
        if (!(bits & 2u)) {
            std::fprintf(stderr, "Error: HAIR file missing points section.\n");
            return out;
        }
        if (!(bits & 1u)) r.get(def);
        r.at += 4u * (128u - (def ? 3u : 4u));
        seg.resize(n);
        for (unsigned i = 0; i < n; ++i) {
            unsigned short v = static_cast<unsigned short>(def);
            if (def == 0) r.get(v);
            seg[i] = static_cast<unsigned>(v) + 1u;
        }
    } else {
        r.at = 0;
        r.get(n);
        r.get(count);
        seg.resize(n);
        for (unsigned i = 0; i < n; ++i) {
            unsigned short v = 0;
            r.get(v);
            seg[i] = v;
        }
    }
    std::vector<float> pts(static_cast<size_t>(count) * 3, 0.f);
    for (float& v : pts) r.get(v);
    size_t at = 0;
    for (unsigned i = 0; i < n; ++i) {
        if ((at + seg[i]) * 3 > pts.size()) break;  // the file's points run out
        out.emplace_back(pts.begin() + at * 3, pts.begin() + (at + seg[i]) * 3);
        at += seg[i];
    }
    return out;
}

}  // namespace

PYBIND11_MODULE(_core, m) {
    m.doc() = "strns2cards_amd native core (gfx950 HIP)";
#ifndef S2C_BUILD_ID
#define S2C_BUILD_ID "unknown"
#endif
    // hash of the sources this binding was built from (build.py:source_hash)
    m.attr("build_id") = S2C_BUILD_ID;

    m.def("print_version", []() {
        const int v = s2c_version();
        py::print("pystrns2cards version " + std::to_string(v / 10000) + "." +
                  std::to_string((v / 100) % 100) + "." + std::to_string(v % 100) +
                  " (strns2cards_amd, gfx950 HIP)");
    });

    // pystrns2cards.cpp:245-281
    m.def("compute_triangle_rects", [](int H, int W, Tensor pos, Tensor tri, Tensor rects,
                                       Tensor prefix) -> int {
        const char* fn = "compute_triangle_rects";
        need(pos, fn, "pos", at::kFloat, {-1, -1, 4});
        need(tri, fn, "tri", at::kInt, {-1, 3});
        need(rects, fn, "triangle_rects", at::kInt, {-1, 4});
        need(prefix, fn, "frag_prefix_sum", at::kInt, {-1});
        same_device(fn, {&pos, &tri, &rects, &prefix});
        const int B = I(pos.size(0), fn), V = I(pos.size(1), fn), T = I(tri.size(0), fn);
        if (rects.size(0) != static_cast<int64_t>(B) * T)
            throw std::runtime_error("`triangle_rects` must have shape [B * T, 4].");
        if (prefix.size(0) != static_cast<int64_t>(B) * T)
            throw std::runtime_error("`frag_prefix_sum` must have shape [B * T].");
        DeviceGuard g(pos.device());
        int n = 0;
        {
            py::gil_scoped_release nogil;
            check(s2c_compute_triangle_rects(H, W, B, V, P<float>(pos), T, P<int>(tri),
                                             P<int>(rects), P<int>(prefix), &n, stream_of(pos)),
                  fn);
        }
        return n;
    });

    // pystrns2cards.cpp:283-326
    m.def("compute_fragments", [](int H, int W, Tensor pos, Tensor tri, Tensor prefix,
                                  Tensor rects, Tensor frag_pix, Tensor frag_attrs) {
        const char* fn = "compute_fragments";
        need(pos, fn, "pos", at::kFloat, {-1, -1, 4});
        need(tri, fn, "tri", at::kInt, {-1, 3});
        need(prefix, fn, "frag_prefix_sum", at::kInt, {-1});
        need(rects, fn, "triangle_rects", at::kInt, {-1, 4});
        need(frag_pix, fn, "frag_pix", at::kInt, {-1, 3});
        need(frag_attrs, fn, "frag_attrs", at::kFloat, {-1, 4});
        same_device(fn, {&pos, &tri, &prefix, &rects, &frag_pix, &frag_attrs});
        const int B = I(pos.size(0), fn), V = I(pos.size(1), fn), T = I(tri.size(0), fn);
        const int num_tris = I(static_cast<int64_t>(B) * T, fn);
        const int N = I(frag_pix.size(0), fn);
        if (frag_attrs.size(0) != N || prefix.size(0) != num_tris)
            throw std::runtime_error("Mismatch in fragment count or prefix sum length.");
        if (rects.size(0) != num_tris)
            throw std::runtime_error("`triangle_rects` must have shape [B * T, 4].");
        DeviceGuard g(pos.device());
        check(s2c_compute_fragments(H, W, V, P<float>(pos), T, P<int>(tri), num_tris, N,
                                    P<int>(prefix), P<int>(rects), P<int>(frag_pix),
                                    P<float>(frag_attrs), stream_of(pos)),
              fn);
    });

    // pystrns2cards.cpp:328-356
    m.def("depth_test", [](int H, int W, Tensor frag_pix, Tensor frag_attrs, Tensor rast_out) {
        const char* fn = "depth_test";
        need(frag_pix, fn, "frag_pix", at::kInt, {-1, 3});
        need(frag_attrs, fn, "frag_attrs", at::kFloat, {-1, 4});
        need(rast_out, fn, "rast_out", at::kFloat, {-1, -1, -1, 4});
        same_device(fn, {&frag_pix, &frag_attrs, &rast_out});
        const int N = I(frag_pix.size(0), fn);
        if (frag_attrs.size(0) != N)
            throw std::runtime_error("frag_attrs must match frag_pix in size.");
        if (rast_out.size(1) != H || rast_out.size(2) != W)
            throw std::runtime_error("rast_out must have shape (B, H, W, 4).");
        DeviceGuard g(rast_out.device());
        check(s2c_depth_test(I(rast_out.size(0), fn), H, W, N, P<int>(frag_pix),
                             P<float>(frag_attrs), P<float>(rast_out), stream_of(rast_out)),
              fn);
    });

    // pystrns2cards.cpp:358-392
    m.def("filter_valid_fragments", [](Tensor fp, Tensor fa, Tensor fp_out, Tensor fa_out) -> int {
        const char* fn = "filter_valid_fragments";
        need(fp, fn, "frag_pix", at::kInt, {-1, 3});
        need(fa, fn, "frag_attrs", at::kFloat, {-1, 4});
        need(fp_out, fn, "frag_pix_out", at::kInt, {-1, 3});
        need(fa_out, fn, "frag_attrs_out", at::kFloat, {-1, 4});
        same_device(fn, {&fp, &fa, &fp_out, &fa_out});
        const int N = I(fp.size(0), fn);
        if (fa.size(0) != N || fp_out.size(0) != N || fa_out.size(0) != N)
            throw std::runtime_error("Mismatch in fragment counts between input/output tensors.");
        DeviceGuard g(fp.device());
        int n = 0;
        {
            py::gil_scoped_release nogil;
            check(s2c_filter_valid_fragments(N, P<int>(fp), P<float>(fa), P<int>(fp_out),
                                             P<float>(fa_out), &n, stream_of(fp)),
                  fn);
        }
        return n;
    });

    // pystrns2cards.cpp:394-415
    m.def("interpolate_triangle_attributes", [](Tensor rast, Tensor attr, Tensor tri, Tensor image) {
        const char* fn = "interpolate_triangle_attributes";
        need(rast, fn, "rast", at::kFloat, {-1, -1, -1, 4});
        need(attr, fn, "attr", at::kFloat, {-1, -1});
        need(tri, fn, "tri", at::kInt, {-1, 3});
        need(image, fn, "image", at::kFloat, {-1, -1, -1, -1});
        same_device(fn, {&rast, &attr, &tri, &image});
        const int B = I(rast.size(0), fn), H = I(rast.size(1), fn), W = I(rast.size(2), fn);
        const int C = I(attr.size(1), fn);
        if (image.size(0) != B || image.size(1) != H || image.size(2) != W || image.size(3) != C)
            throw std::runtime_error("image must have shape [B, H, W, C]");
        DeviceGuard g(rast.device());
        check(s2c_interpolate_triangle_attributes(B, H, W, C, P<float>(rast), P<float>(attr),
                                                  P<int>(tri), P<float>(image), stream_of(rast)),
              fn);
    });

    // pystrns2cards.cpp:417-438
    m.def("backward_interpolate_triangle_attributes",
          [](Tensor rast, Tensor d_image, Tensor tri, Tensor d_attr) {
              const char* fn = "backward_interpolate_triangle_attributes";
              need(rast, fn, "rast", at::kFloat, {-1, -1, -1, 4});
              need(d_image, fn, "d_image", at::kFloat, {-1, -1, -1, -1});
              need(tri, fn, "tri", at::kInt, {-1, 3});
              need(d_attr, fn, "d_attr", at::kFloat, {-1, -1});
              same_device(fn, {&rast, &d_image, &tri, &d_attr});
              const int B = I(rast.size(0), fn), H = I(rast.size(1), fn), W = I(rast.size(2), fn);
              const int C = I(d_attr.size(1), fn);
              if (d_image.size(0) != B || d_image.size(1) != H || d_image.size(2) != W ||
                  d_image.size(3) != C)
                  throw std::runtime_error("`d_image` must have shape [B, H, W, C] matching rast");
              DeviceGuard g(rast.device());
              check(s2c_backward_interpolate_triangle_attributes(
                        B, H, W, C, P<float>(rast), P<float>(d_image), P<int>(tri),
                        P<float>(d_attr), stream_of(rast)),
                    fn);
          });

    // pystrns2cards.cpp:440-478
    m.def("cluster_mask_from_fragments",
          [](Tensor fp, Tensor fa, Tensor prim2cluster, Tensor bitset) {
              const char* fn = "cluster_mask_from_fragments";
              need(fp, fn, "frag_pix", at::kInt, {-1, 3});
              need(fa, fn, "frag_attrs", at::kFloat, {-1, 4});
              need(prim2cluster, fn, "prim2cluster", at::kInt, {-1});
              need(bitset, fn, "bitset", at::kInt, {-1, -1, -1, -1});
              same_device(fn, {&fp, &fa, &prim2cluster, &bitset});
              const int N = I(fp.size(0), fn);
              if (fa.size(0) != N)
                  throw std::runtime_error(
                      "frag_pix and frag_attrs must have the same number of fragments.");
              DeviceGuard g(bitset.device());
              check(s2c_cluster_mask_from_fragments(N, P<int>(fp), P<float>(fa),
                                                    P<int>(prim2cluster), I(bitset.size(1), fn),
                                                    I(bitset.size(2), fn), I(bitset.size(3), fn),
                                                    P<int>(bitset), stream_of(bitset)),
                    fn);
          });

    // rasterize.py:170-214 (zeros, then cluster_mask_from_fragments) as one
    // binned pass writing every word of `bitset`
    m.def("fragment_cluster_mask", [](Tensor fp, Tensor fa, Tensor prim2cluster, Tensor bitset) {
        const char* fn = "fragment_cluster_mask";
        need(fp, fn, "frag_pix", at::kInt, {-1, 3});
        need(fa, fn, "frag_attrs", at::kFloat, {-1, 4});
        need(prim2cluster, fn, "prim2cluster", at::kInt, {-1});
        need(bitset, fn, "bitset", at::kInt, {-1, -1, -1, -1});
        same_device(fn, {&fp, &fa, &prim2cluster, &bitset});
        const int N = I(fp.size(0), fn);
        if (fa.size(0) != N)
            throw std::runtime_error(
                "frag_pix and frag_attrs must have the same number of fragments.");
        DeviceGuard g(bitset.device());
        check(s2c_fragment_cluster_mask(I(bitset.size(0), fn), I(bitset.size(1), fn),
                                        I(bitset.size(2), fn), N, P<int>(fp), P<float>(fa),
                                        P<int>(prim2cluster), I(bitset.size(3), fn),
                                        P<int>(bitset), stream_of(bitset)),
              fn);
    });

    m.def("max_frag_view", [](Tensor fp) -> int {
        const char* fn = "max_frag_view";
        need(fp, fn, "frag_pix", at::kInt, {-1, 3});
        if (fp.size(0) == 0) throw std::runtime_error("max_frag_view: no fragments");
        DeviceGuard g(fp.device());
        int m = 0;
        {
            py::gil_scoped_release nogil;
            check(s2c_max_frag_view(I(fp.size(0), fn), P<int>(fp), &m, stream_of(fp)), fn);
        }
        return m;
    });

    // pystrns2cards.cpp:480-505
    m.def("accumulate_bitset_rgb", [](Tensor bitset, Tensor rgb, Tensor out) {
        const char* fn = "accumulate_bitset_rgb";
        need(bitset, fn, "bitset", at::kInt, {-1, -1, -1, -1});
        need(rgb, fn, "cluster_rgb", at::kFloat, {-1, 3});
        need(out, fn, "accum_color", at::kFloat, {-1, -1, -1, 3});
        same_device(fn, {&bitset, &rgb, &out});
        const int B = I(bitset.size(0), fn), H = I(bitset.size(1), fn), W = I(bitset.size(2), fn);
        const int S = I(bitset.size(3), fn);
        if (static_cast<int64_t>(S) * 32 < rgb.size(0))
            throw std::runtime_error("Too few slots for the number of clusters.");
        if (out.size(0) != B || out.size(1) != H || out.size(2) != W)
            throw std::runtime_error("accum_color must have shape [B, H, W, 3]");
        DeviceGuard g(bitset.device());
        check(s2c_accumulate_bitset_rgb(B, H, W, S, P<int>(bitset), P<float>(rgb), P<float>(out),
                                        stream_of(bitset)),
              fn);
    });

    // pystrns2cards.cpp:507-529
    m.def("popcount_bitset", [](Tensor bitset, Tensor out) {
        const char* fn = "popcount_bitset";
        need(bitset, fn, "bitset", at::kInt, {-1, -1, -1, -1});
        need(out, fn, "count_image", at::kFloat, {-1, -1, -1});
        same_device(fn, {&bitset, &out});
        const int B = I(bitset.size(0), fn), H = I(bitset.size(1), fn), W = I(bitset.size(2), fn);
        if (out.size(0) != B || out.size(1) != H || out.size(2) != W)
            throw std::runtime_error("count_image must have shape [B, H, W]");
        DeviceGuard g(bitset.device());
        check(s2c_popcount_bitset(B, H, W, I(bitset.size(3), fn), P<int>(bitset), P<float>(out),
                                  stream_of(bitset)),
              fn);
    });

    // pystrns2cards.cpp:531-565
    m.def("dda_compute_span", [](Tensor edge_ids, int H, int W, Tensor pos, Tensor edges,
                                 Tensor prefix, Tensor slopes, Tensor spans) -> int {
        const char* fn = "dda_compute_span";
        need(edge_ids, fn, "edge_ids", at::kInt, {-1});
        need(pos, fn, "pos", at::kFloat, {-1, -1, 4});
        need(edges, fn, "edges", at::kInt, {-1, 2});
        need(prefix, fn, "frag_prefix_sum", at::kInt, {-1});
        need(slopes, fn, "frag_slopes", at::kFloat, {-1, 2});
        need(spans, fn, "frag_spans", at::kFloat, {-1, 4});
        same_device(fn, {&edge_ids, &pos, &edges, &prefix, &slopes, &spans});
        const int P_ = I(edge_ids.size(0), fn);
        if (prefix.size(0) != P_ || slopes.size(0) != P_ || spans.size(0) != P_)
            throw std::runtime_error(
                "Invalid output tensor shapes for frag_prefix_sum, frag_slopes, or frag_spans.");
        DeviceGuard g(pos.device());
        int n = 0;
        {
            py::gil_scoped_release nogil;
            check(s2c_dda_compute_span(P_, P<int>(edge_ids), H, W, I(pos.size(1), fn),
                                       P<float>(pos), I(edges.size(0), fn), P<int>(edges),
                                       P<int>(prefix), P<float>(slopes), P<float>(spans), &n,
                                       stream_of(pos)),
                  fn);
        }
        return n;
    });

    // pystrns2cards.cpp:567-611
    m.def("dda_compute_fragments", [](Tensor prefix, Tensor edge_ids, Tensor slopes, Tensor spans,
                                      int H, int W, Tensor pos, Tensor edges, Tensor frag_pix,
                                      Tensor frag_attrs) {
        const char* fn = "dda_compute_fragments";
        need(prefix, fn, "frag_prefix_sum", at::kInt, {-1});
        need(edge_ids, fn, "edge_ids", at::kInt, {-1});
        need(slopes, fn, "frag_slopes", at::kFloat, {-1, 2});
        need(spans, fn, "frag_spans", at::kFloat, {-1, 4});
        need(pos, fn, "pos", at::kFloat, {-1, -1, 4});
        need(edges, fn, "edges", at::kInt, {-1, 2});
        need(frag_pix, fn, "frag_pix", at::kInt, {-1, 3});
        need(frag_attrs, fn, "frag_attrs", at::kFloat, {-1, 4});
        same_device(fn, {&prefix, &edge_ids, &slopes, &spans, &pos, &edges, &frag_pix, &frag_attrs});
        const int N = I(frag_pix.size(0), fn), P_ = I(edge_ids.size(0), fn);
        if (prefix.size(0) != P_ || slopes.size(0) != P_ || spans.size(0) != P_ ||
            frag_attrs.size(0) != N)
            throw std::runtime_error("Invalid shape for input or output arrays.");
        DeviceGuard g(pos.device());
        check(s2c_dda_compute_fragments(P_, N, P<int>(prefix), P<int>(edge_ids), P<float>(slopes),
                                        P<float>(spans), H, W, I(pos.size(1), fn), P<float>(pos),
                                        I(edges.size(0), fn), P<int>(edges), P<int>(frag_pix),
                                        P<float>(frag_attrs), stream_of(pos)),
              fn);
    });

    // dda.py:5-76 in the space of surviving fragments: per-primitive counts
    // (scanned into valid_prefix), then the fragments themselves.  edge_ids
    // None = every edge of every view.
    m.def("dda_count_fragments", [](std::optional<Tensor> edge_ids, int H, int W, Tensor pos,
                                    Tensor edges, Tensor valid_prefix) -> int {
        const char* fn = "dda_count_fragments";
        need(pos, fn, "pos", at::kFloat, {-1, -1, 4});
        need(edges, fn, "edges", at::kInt, {-1, 2});
        need(valid_prefix, fn, "valid_prefix", at::kInt, {-1});
        same_device(fn, {&pos, &edges, &valid_prefix});
        const int P_ = I(valid_prefix.size(0), fn);
        if (edge_ids) {
            need(*edge_ids, fn, "edge_ids", at::kInt, {P_});
            same_device(fn, {&pos, &*edge_ids});
        } else if (P_ != pos.size(0) * edges.size(0)) {
            throw std::runtime_error(std::string(fn) + ": valid_prefix must have B * E entries");
        }
        DeviceGuard g(pos.device());
        int n = 0;
        {
            py::gil_scoped_release nogil;
            check(s2c_dda_count_fragments(P_, edge_ids ? P<int>(*edge_ids) : nullptr, H, W,
                                          I(pos.size(1), fn), P<float>(pos), I(edges.size(0), fn),
                                          P<int>(edges), P<int>(valid_prefix), &n, stream_of(pos)),
                  fn);
        }
        return n;
    });

    m.def("dda_emit_fragments", [](Tensor valid_prefix, std::optional<Tensor> edge_ids, int H,
                                   int W, Tensor pos, Tensor edges, Tensor frag_pix,
                                   Tensor frag_attrs) {
        const char* fn = "dda_emit_fragments";
        need(valid_prefix, fn, "valid_prefix", at::kInt, {-1});
        need(pos, fn, "pos", at::kFloat, {-1, -1, 4});
        need(edges, fn, "edges", at::kInt, {-1, 2});
        need(frag_pix, fn, "frag_pix", at::kInt, {-1, 3});
        need(frag_attrs, fn, "frag_attrs", at::kFloat, {frag_pix.size(0), 4});
        same_device(fn, {&valid_prefix, &pos, &edges, &frag_pix, &frag_attrs});
        const int P_ = I(valid_prefix.size(0), fn);
        if (edge_ids) {
            need(*edge_ids, fn, "edge_ids", at::kInt, {P_});
            same_device(fn, {&pos, &*edge_ids});
        } else if (P_ != pos.size(0) * edges.size(0)) {
            throw std::runtime_error(std::string(fn) + ": valid_prefix must have B * E entries");
        }
        DeviceGuard g(pos.device());
        check(s2c_dda_emit_fragments(P_, I(frag_pix.size(0), fn), P<int>(valid_prefix),
                                     edge_ids ? P<int>(*edge_ids) : nullptr, H, W,
                                     I(pos.size(1), fn), P<float>(pos), I(edges.size(0), fn),
                                     P<int>(edges), P<int>(frag_pix), P<float>(frag_attrs),
                                     stream_of(pos)),
              fn);
    });

    // pystrns2cards.cpp:613-647
    m.def("dda_interpolate_attributes", [](Tensor rast, Tensor attr, Tensor edges, Tensor image) {
        const char* fn = "dda_interpolate_attributes";
        need(rast, fn, "rast_dda", at::kFloat, {-1, -1, -1, 4});
        need(attr, fn, "attr", at::kFloat, {-1, -1});
        need(edges, fn, "edges", at::kInt, {-1, 2});
        need(image, fn, "image", at::kFloat, {-1, -1, -1, -1});
        same_device(fn, {&rast, &attr, &edges, &image});
        const int B = I(rast.size(0), fn), H = I(rast.size(1), fn), W = I(rast.size(2), fn);
        const int C = I(attr.size(1), fn);
        if (image.size(0) != B || image.size(1) != H || image.size(2) != W || image.size(3) != C)
            throw std::runtime_error("image must have shape [B, H, W, C]");
        DeviceGuard g(rast.device());
        check(s2c_dda_interpolate_attributes(B, H, W, C, P<float>(rast), P<float>(attr),
                                             P<int>(edges), P<float>(image), stream_of(rast)),
              fn);
    });

    // pystrns2cards.cpp:649-680
    m.def("backward_dda_interpolate_attributes",
          [](Tensor rast, Tensor d_image, Tensor edges, Tensor d_attr) {
              const char* fn = "backward_dda_interpolate_attributes";
              need(rast, fn, "rast_dda", at::kFloat, {-1, -1, -1, 4});
              need(d_image, fn, "d_image", at::kFloat, {-1, -1, -1, -1});
              need(edges, fn, "edges", at::kInt, {-1, 2});
              need(d_attr, fn, "d_attr", at::kFloat, {-1, -1});
              same_device(fn, {&rast, &d_image, &edges, &d_attr});
              const int B = I(rast.size(0), fn), H = I(rast.size(1), fn), W = I(rast.size(2), fn);
              const int C = I(d_image.size(3), fn);
              if (d_image.size(0) != B || d_image.size(1) != H || d_image.size(2) != W)
                  throw std::runtime_error(
                      "`d_image` must have shape [B, H, W, C] matching rast_dda");
              if (d_attr.size(1) != C)
                  throw std::runtime_error(
                      "`d_attr` must have shape [V, C] where C matches d_image.shape(3)");
              DeviceGuard g(rast.device());
              check(s2c_backward_dda_interpolate_attributes(B, H, W, P<float>(rast), C,
                                                            P<float>(d_image), P<int>(edges),
                                                            P<float>(d_attr), stream_of(rast)),
                    fn);
          });

    // pystrns2cards.cpp:682-715
    m.def("mark_discontinuity_edges", [](int H, int W, Tensor tri, Tensor edges, Tensor edge2tri,
                                         Tensor pos, Tensor prim_ids, Tensor normals) {
        const char* fn = "mark_discontinuity_edges";
        need(tri, fn, "tri", at::kInt, {-1, 3});
        need(edges, fn, "edges", at::kInt, {-1, 2});
        need(edge2tri, fn, "edge2tri", at::kInt, {-1, 2});
        need(pos, fn, "pos", at::kFloat, {-1, -1, 4});
        need(prim_ids, fn, "prim_ids", at::kInt, {-1});
        need(normals, fn, "normals", at::kFloat, {-1, 2});
        same_device(fn, {&tri, &edges, &edge2tri, &pos, &prim_ids, &normals});
        const int B = I(pos.size(0), fn), V = I(pos.size(1), fn), E = I(edges.size(0), fn);
        if (edge2tri.size(0) != E)
            throw std::runtime_error(
                "mark_discontinuity_edges: edge2tri must have the same number of edges.");
        if (prim_ids.size(0) != static_cast<int64_t>(B) * E)
            throw std::runtime_error("mark_discontinuity_edges: prim_ids must have shape (B * E).");
        if (normals.size(0) != static_cast<int64_t>(B) * E)
            throw std::runtime_error(
                "mark_discontinuity_edges: normals must have shape (B * E, 2).");
        DeviceGuard g(pos.device());
        check(s2c_mark_discontinuity_edges(H, W, B, V, P<float>(pos), P<int>(tri), E,
                                           P<int>(edges), P<int>(edge2tri), P<int>(prim_ids),
                                           P<float>(normals), stream_of(pos)),
              fn);
    });

    // pystrns2cards.cpp:717-772
    m.def("backward_antialiased_cluster_bitset",
          [](Tensor fp, Tensor fa, Tensor bitset, Tensor target, Tensor pos, Tensor edges,
             Tensor normals, Tensor edge2cluster, Tensor d_pos, float kernel_radius, float rho) {
              const char* fn = "backward_antialiased_cluster_bitset";
              need(fp, fn, "frag_pix", at::kInt, {-1, 3});
              need(fa, fn, "frag_attrs_dda", at::kFloat, {-1, 4});
              need(bitset, fn, "bitset", at::kInt, {-1, -1, -1, -1});
              need(target, fn, "target_bitset", at::kInt, {-1, -1, -1, -1});
              need(pos, fn, "pos", at::kFloat, {-1, -1, 4});
              need(edges, fn, "edges", at::kInt, {-1, 2});
              need(normals, fn, "normals", at::kFloat, {-1, 2});
              need(edge2cluster, fn, "edge2cluster", at::kInt, {-1});
              need(d_pos, fn, "d_pos", at::kFloat, {-1, -1, 4});
              same_device(fn, {&fp, &fa, &bitset, &target, &pos, &edges, &normals, &edge2cluster,
                               &d_pos});
              const int N = I(fp.size(0), fn);
              const int B = I(bitset.size(0), fn), H = I(bitset.size(1), fn);
              const int W = I(bitset.size(2), fn), S = I(bitset.size(3), fn);
              const int V = I(pos.size(1), fn), E = I(edges.size(0), fn);
              if (fa.size(0) != N)
                  throw std::runtime_error(
                      "backward_antialiased_bitset: frag_pix and frag_attrs_dda must match in "
                      "number of fragments.");
              if (target.size(0) != B || target.size(1) != H || target.size(2) != W ||
                  target.size(3) != S)
                  throw std::runtime_error(
                      "backward_antialiased_bitset: target_bitset must match bitset shape.");
              if (d_pos.size(0) != B || d_pos.size(1) != V)
                  throw std::runtime_error(
                      "backward_antialiased_bitset: d_pos must match pos shape.");
              DeviceGuard g(pos.device());
              check(s2c_backward_antialiased_cluster_bitset_views(
                        B, N, P<int>(fp), P<float>(fa), H, W, S, P<int>(bitset), P<int>(target), V,
                        P<float>(pos), E, P<int>(edges), P<float>(normals), P<int>(edge2cluster),
                        P<float>(d_pos), kernel_radius, rho, stream_of(pos)),
                    fn);
          });

    // The same gradient from the marked edges themselves (prim_ids of
    // mark_discontinuity_edges before compaction: b * E + e or -1), with the
    // DDA walk inside the kernel: no fragment list and no host sync.
    m.def("backward_antialiased_cluster_bitset_edges",
          [](Tensor prim_ids, Tensor normals, Tensor bitset, Tensor target, Tensor pos,
             Tensor edges, Tensor edge2cluster, Tensor d_pos, float kernel_radius, float rho) {
              const char* fn = "backward_antialiased_cluster_bitset_edges";
              need(prim_ids, fn, "prim_ids", at::kInt, {-1});
              need(normals, fn, "normals", at::kFloat, {-1, 2});
              need(bitset, fn, "bitset", at::kInt, {-1, -1, -1, -1});
              need(target, fn, "target_bitset", at::kInt, {-1, -1, -1, -1});
              need(pos, fn, "pos", at::kFloat, {-1, -1, 4});
              need(edges, fn, "edges", at::kInt, {-1, 2});
              need(edge2cluster, fn, "edge2cluster", at::kInt, {-1});
              need(d_pos, fn, "d_pos", at::kFloat, {-1, -1, 4});
              same_device(fn, {&prim_ids, &normals, &bitset, &target, &pos, &edges,
                               &edge2cluster, &d_pos});
              const int B = I(bitset.size(0), fn), H = I(bitset.size(1), fn);
              const int W = I(bitset.size(2), fn), S = I(bitset.size(3), fn);
              const int V = I(pos.size(1), fn), E = I(edges.size(0), fn);
              if (pos.size(0) != B)
                  throw std::runtime_error(std::string(fn) +
                                           ": pos must hold one view per bitset view.");
              if (prim_ids.size(0) != pos.size(0) * edges.size(0) ||
                  normals.size(0) != prim_ids.size(0))
                  throw std::runtime_error(std::string(fn) +
                                           ": prim_ids and normals must have B * E rows.");
              if (edge2cluster.size(0) != E)
                  throw std::runtime_error(std::string(fn) +
                                           ": edge2cluster must have one entry per edge.");
              if (target.sizes() != bitset.sizes())
                  throw std::runtime_error(std::string(fn) +
                                           ": target_bitset must match bitset shape.");
              if (d_pos.sizes() != pos.sizes())
                  throw std::runtime_error(std::string(fn) + ": d_pos must match pos shape.");
              DeviceGuard g(pos.device());
              check(s2c_backward_antialiased_cluster_bitset_edges(
                        B, H, W, S, P<int>(bitset), P<int>(target), V, P<float>(pos), E,
                        P<int>(edges), P<int>(prim_ids), P<float>(normals), P<int>(edge2cluster),
                        P<float>(d_pos), kernel_radius, rho, stream_of(pos)),
                    fn);
          });

    // optimizer.py:33-68 (UniformAdam.step) for a (n, dim) parameter.
    m.def("uniform_adam_step", [](Tensor param, Tensor grad, Tensor g1, Tensor g2, double lr,
                                  double b1, double b2, int step) {
        const char* fn = "uniform_adam_step";
        need(param, fn, "param", at::kFloat, {-1, -1});
        const int64_t n = param.size(0), dim = param.size(1);
        need(grad, fn, "grad", at::kFloat, {n, dim});
        need(g1, fn, "g1", at::kFloat, {n, dim});
        need(g2, fn, "g2", at::kFloat, {n});
        same_device(fn, {&param, &grad, &g1, &g2});
        DeviceGuard g(param.device());
        check(s2c_uniform_adam_step(I(n, fn), I(dim, fn), lr, b1, b2, step, P<float>(param),
                                    P<float>(grad), P<float>(g1), P<float>(g2),
                                    stream_of(param)),
              fn);
    });

    // pystrns2cards.cpp:774-786
    m.def("compact_valid_ints", [](Tensor in, Tensor out) -> int {
        const char* fn = "compact_valid_ints";
        need(in, fn, "input", at::kInt, {-1});
        need(out, fn, "output", at::kInt, {-1});
        same_device(fn, {&in, &out});
        if (in.size(0) != out.size(0))
            throw std::runtime_error(
                "compact_valid_ints: input and output must have the same size.");
        DeviceGuard g(in.device());
        int n = 0;
        {
            py::gil_scoped_release nogil;
            check(s2c_compact_valid_ints(I(in.size(0), fn), P<int>(in), P<int>(out), &n,
                                         stream_of(in)),
                  fn);
        }
        return n;
    });

    // pystrns2cards.cpp:841-871
    m.def("smooth_strands", [](Tensor strands, Tensor arclength, Tensor out, float beta) {
        const char* fn = "smooth_strands";
        need(strands, fn, "strands", at::kFloat, {-1, -1, 3});
        need(arclength, fn, "arclength", at::kFloat, {-1, -1});
        need(out, fn, "smoothed_strands", at::kFloat, {-1, -1, 3});
        same_device(fn, {&strands, &arclength, &out});
        const int N = I(strands.size(0), fn), Pn = I(strands.size(1), fn);
        if (arclength.size(0) != N || arclength.size(1) != Pn)
            throw std::runtime_error(
                "smooth_strands: arclength must match (num_strands, num_points).");
        if (out.size(0) != N || out.size(1) != Pn)
            throw std::runtime_error(
                "smooth_strands: smoothed_strands must match (num_strands, num_points, 3).");
        DeviceGuard g(strands.device());
        check(s2c_smooth_strands(N, Pn, P<float>(strands), P<float>(arclength), P<float>(out),
                                 beta, stream_of(strands)),
              fn);
    });

    // pystrns2cards.cpp:873-923
    m.def("skinning", [](Tensor strands, Tensor arclength, Tensor handle, Tensor canonical,
                         Tensor R, Tensor out, float beta) {
        const char* fn = "skinning";
        need(strands, fn, "strands", at::kFloat, {-1, -1, 3});
        need(arclength, fn, "arclength", at::kFloat, {-1, -1});
        need(handle, fn, "handle_strands", at::kFloat, {-1, -1, 3});
        need(canonical, fn, "canonical_handles", at::kFloat, {-1, -1, 3});
        need(R, fn, "R", at::kFloat, {-1, -1, 3, 3});
        need(out, fn, "skinned_strands", at::kFloat, {-1, -1, 3});
        same_device(fn, {&strands, &arclength, &handle, &canonical, &R, &out});
        const int N = I(strands.size(0), fn), Pn = I(strands.size(1), fn);
        auto match = [&](const Tensor& t) { return t.size(0) == N && t.size(1) == Pn; };
        if (!match(arclength))
            throw std::runtime_error("skinning: arclength must match (num_strands, num_points).");
        if (!match(handle))
            throw std::runtime_error(
                "skinning: handle_strands must match (num_strands, num_points, 3).");
        if (!match(canonical))
            throw std::runtime_error(
                "skinning: canonical_handles must match (num_strands, num_points, 3).");
        if (!match(R))
            throw std::runtime_error("skinning: R must match (num_strands, num_points, 3, 3).");
        if (!match(out))
            throw std::runtime_error(
                "skinning: skinned_strands must match (num_strands, num_points, 3).");
        DeviceGuard g(strands.device());
        check(s2c_skinning(N, Pn, P<float>(strands), P<float>(arclength), P<float>(handle),
                           P<float>(canonical), P<float>(R), P<float>(out), beta,
                           stream_of(strands)),
              fn);
    });

    // Fused forms (no single reference binding; see the C ABI header).
    m.def("dda_cluster_mask", [](Tensor pos, Tensor edges, Tensor edge2cluster, Tensor bitset,
                                 int tile_log2) {
        const char* fn = "dda_cluster_mask";
        need(pos, fn, "pos", at::kFloat, {-1, -1, 4});
        need(edges, fn, "edges", at::kInt, {-1, 2});
        need(edge2cluster, fn, "edge2cluster", at::kInt, {-1});
        need(bitset, fn, "bitset", at::kInt, {-1, -1, -1, -1});
        same_device(fn, {&pos, &edges, &edge2cluster, &bitset});
        if (bitset.size(0) != pos.size(0))
            throw std::runtime_error("dda_cluster_mask: bitset must have B = pos.shape[0] views");
        if (edge2cluster.size(0) != edges.size(0))
            throw std::runtime_error("dda_cluster_mask: edge2cluster must have E entries");
        DeviceGuard g(pos.device());
        check(s2c_dda_cluster_mask_tiled(I(pos.size(0), fn), I(bitset.size(1), fn),
                                         I(bitset.size(2), fn), I(pos.size(1), fn), P<float>(pos),
                                         I(edges.size(0), fn), P<int>(edges), P<int>(edge2cluster),
                                         I(bitset.size(3), fn), tile_log2, P<int>(bitset),
                                         stream_of(pos)),
              fn);
    });

    // Batched card fitting (see the C ABI header).
    m.def("fit_targets", [](Tensor res, Tensor bits_off, Tensor strand_off, Tensor num_strands,
                            Tensor strand_clip, Tensor target_bits) {
        const char* fn = "fit_targets";
        need(res, fn, "res", at::kInt, {-1});
        const int64_t J = res.size(0);
        need(bits_off, fn, "bits_off", at::kInt, {J});
        need(strand_off, fn, "strand_off", at::kInt, {J});
        need(num_strands, fn, "num_strands", at::kInt, {J});
        need(strand_clip, fn, "strand_clip", at::kFloat, {-1, -1, 4});
        need(target_bits, fn, "target_bits", at::kInt, {-1});
        same_device(fn, {&res, &bits_off, &strand_off, &num_strands, &strand_clip, &target_bits});
        DeviceGuard g(res.device());
        check(s2c_fit_targets(I(J, fn), P<int>(res), P<int>(bits_off), P<int>(strand_off),
                              P<int>(num_strands), I(strand_clip.size(1), fn),
                              P<float>(strand_clip), P<unsigned>(target_bits), stream_of(res)),
              fn);
    });

    m.def("fit_admit", [](Tensor records, Tensor frecords, Tensor slots, Tensor res, Tensor quads,
                          Tensor bits_off, std::optional<Tensor> limit, Tensor mvp, Tensor steps,
                          Tensor verts, Tensor adam_m, Tensor adam_v) {
        const char* fn = "fit_admit";
        need(records, fn, "records", at::kInt, {-1, 8});
        const int64_t n = records.size(0);
        need(frecords, fn, "frecords", at::kFloat, {n, -1});
        need(slots, fn, "slots", at::kInt, {-1});
        const int64_t J = slots.size(0);
        need(res, fn, "res", at::kInt, {J});
        need(quads, fn, "quads", at::kInt, {J});
        need(bits_off, fn, "bits_off", at::kInt, {J});
        need(mvp, fn, "mvp", at::kFloat, {J, 4, 4});
        need(steps, fn, "steps", at::kInt, {J});
        need(verts, fn, "verts", at::kFloat, {-1, 3});
        need(adam_m, fn, "adam_m", at::kFloat, {verts.size(0), 3});
        need(adam_v, fn, "adam_v", at::kFloat, {verts.size(0)});
        same_device(fn, {&records, &frecords, &slots, &res, &quads, &bits_off, &mvp, &steps,
                         &verts, &adam_m, &adam_v});
        if (limit) {
            need(*limit, fn, "limit", at::kInt, {J});
            same_device(fn, {&res, &*limit});
        }
        if (J == 0 || verts.size(0) % J != 0)
            throw std::runtime_error("fit_admit: verts must hold max_verts rows per slot");
        if (frecords.size(1) != 16 + 3 * (verts.size(0) / J) && n > 0)
            throw std::runtime_error("fit_admit: frecords rows must hold 16 + 3 * max_verts floats");
        DeviceGuard g(res.device());
        check(s2c_fit_admit(I(n, fn), P<int>(records), P<float>(frecords), I(J, fn),
                            I(verts.size(0) / J, fn), P<int>(slots), P<int>(res), P<int>(quads),
                            P<int>(bits_off), limit ? P<int>(*limit) : nullptr, P<float>(mvp),
                            P<int>(steps), P<float>(verts), P<float>(adam_m), P<float>(adam_v),
                            stream_of(res)),
              fn);
    });

    m.def("fit_cards", [](Tensor slots, int iters, Tensor res, Tensor quads, Tensor vert_off,
                          Tensor bits_off, Tensor mvp, Tensor verts, Tensor params, Tensor adam_m,
                          Tensor adam_v, Tensor steps, std::optional<Tensor> limit,
                          Tensor target_bits, std::vector<float> hyper, Tensor loss_out,
                          int chunk) {
        const char* fn = "fit_cards";
        need(slots, fn, "slots", at::kInt, {-1});
        need(res, fn, "res", at::kInt, {-1});
        const int64_t J = res.size(0);
        need(quads, fn, "quads", at::kInt, {J});
        need(vert_off, fn, "vert_off", at::kInt, {J});
        need(bits_off, fn, "bits_off", at::kInt, {J});
        need(mvp, fn, "mvp", at::kFloat, {J, 4, 4});
        need(verts, fn, "verts", at::kFloat, {-1, 3});
        need(params, fn, "params", at::kFloat, {verts.size(0), 3});
        need(adam_m, fn, "adam_m", at::kFloat, {verts.size(0), 3});
        need(adam_v, fn, "adam_v", at::kFloat, {verts.size(0)});
        need(steps, fn, "steps", at::kInt, {J});
        need(target_bits, fn, "target_bits", at::kInt, {-1});
        need(loss_out, fn, "loss_out", at::kFloat, {J});
        same_device(fn, {&slots, &res, &quads, &vert_off, &bits_off, &mvp, &verts, &params,
                         &adam_m, &adam_v, &steps, &target_bits, &loss_out});
        if (limit) {
            need(*limit, fn, "limit", at::kInt, {J});
            same_device(fn, {&res, &*limit});
        }
        if (hyper.size() != 10)
            throw std::runtime_error("fit_cards: hyper must hold 10 values");
        if (iters < 0) throw std::runtime_error("fit_cards: iters must be >= 0");
        DeviceGuard g(res.device());
        check(s2c_fit_cards(I(slots.size(0), fn), P<int>(slots), iters, P<int>(res), P<int>(quads),
                            P<int>(vert_off), P<int>(bits_off), P<float>(mvp), P<float>(verts),
                            P<float>(params), P<float>(adam_m), P<float>(adam_v), P<int>(steps),
                            limit ? P<int>(*limit) : nullptr, P<unsigned>(target_bits),
                            hyper.data(), P<float>(loss_out), chunk, stream_of(res)),
              fn);
    });

    m.def("bishop_frame", [](Tensor strands, Tensor T, Tensor N, Tensor B) {
        const char* fn = "bishop_frame";
        need(strands, fn, "strands", at::kFloat, {-1, -1, 3});
        need(T, fn, "T", at::kFloat, {strands.size(0), strands.size(1), 3});
        need(N, fn, "N", at::kFloat, {strands.size(0), strands.size(1), 3});
        need(B, fn, "B", at::kFloat, {strands.size(0), strands.size(1), 3});
        same_device(fn, {&strands, &T, &N, &B});
        DeviceGuard g(strands.device());
        check(s2c_bishop_frame(I(strands.size(0), fn), I(strands.size(1), fn), P<float>(strands),
                               P<float>(T), P<float>(N), P<float>(B), stream_of(strands)),
              fn);
    });

    // ---- host-side compat exports (pystrns2cards.cpp:208-243, 788-839) ----

    m.def("build_edges_from_triangles", [](py::array tri_in) {
        // mesh_utils.cpp:15-79: the undirected edges of the triangles, each
        // with its first two triangles (-1 when it has one).  Edges come in
        // first-appearance order over the sides (0,1), (1,2), (2,0) of each
        // triangle (the reference's order comes from a hash map and is
        // implementation-defined).
        if (tri_in.dtype().kind() != 'i' || tri_in.itemsize() != 4)
            throw py::type_error("build_edges_from_triangles: `tri` must be int32");
        if (tri_in.ndim() != 2 || tri_in.shape(1) != 3)
            throw py::type_error("build_edges_from_triangles: `tri` must have shape (T, 3)");
        py::array_t<int32_t, py::array::c_style | py::array::forcecast> tri(tri_in);
        const int64_t T = tri.shape(0);
        const int32_t* t = tri.data();
        std::unordered_map<uint64_t, int> id;
        id.reserve(static_cast<size_t>(T) * 2);
        std::vector<int32_t> edges, e2t;
        for (int64_t k = 0; k < T; ++k) {
            for (int side = 0; side < 3; ++side) {
                int32_t a = t[3 * k + side], b = t[3 * k + (side + 1) % 3];
                if (a > b) std::swap(a, b);
                const uint64_t key = (static_cast<uint64_t>(static_cast<uint32_t>(a)) << 32) |
                                     static_cast<uint32_t>(b);
                auto it = id.find(key);
                if (it == id.end()) {
                    id.emplace(key, static_cast<int>(edges.size() / 2));
                    edges.push_back(a);
                    edges.push_back(b);
                    e2t.push_back(static_cast<int32_t>(k));
                    e2t.push_back(-1);
                } else if (e2t[2 * it->second + 1] < 0) {
                    e2t[2 * it->second + 1] = static_cast<int32_t>(k);
                }
            }
        }
        const py::ssize_t E = static_cast<py::ssize_t>(edges.size() / 2);
        py::array_t<int32_t> out_e({E, static_cast<py::ssize_t>(2)});
        py::array_t<int32_t> out_t({E, static_cast<py::ssize_t>(2)});
        if (E) {
            std::memcpy(out_e.mutable_data(), edges.data(), sizeof(int32_t) * edges.size());
            std::memcpy(out_t.mutable_data(), e2t.data(), sizeof(int32_t) * e2t.size());
        }
        return py::make_tuple(out_e, out_t);
    }, "Build unique edge list and triangle mapping from face indices");

    m.def("load_strands", [](const std::string& path) {
        // (num_samples (N, 1) int32, points (sum, 3) float32) of the strands
        // with at least 2 points (hair_loader.cpp:220-258), by extension
        auto ends = [&](const char* sfx) {
            const size_t n = std::strlen(sfx);
            return path.size() >= n && path.compare(path.size() - n, n, sfx) == 0;
        };
        std::vector<Strand> strands;
        if (!ends(".bin") && !ends(".data") && !ends(".hair"))
            throw std::invalid_argument("load_strands: unknown file extension.");
        Bytes r;
        if (read_file(path, r)) {
            strands = ends(".bin") ? strands_bin(r) : ends(".data") ? strands_usc(r)
                                                                    : strands_hair(r);
        }
        std::vector<int32_t> ns;
        size_t npts = 0;
        for (const Strand& s : strands)
            if (s.size() >= 6) {
                ns.push_back(static_cast<int32_t>(s.size() / 3));
                npts += s.size() / 3;
            }
        py::array_t<int32_t> num({static_cast<py::ssize_t>(ns.size()), static_cast<py::ssize_t>(1)});
        py::array_t<float> pts({static_cast<py::ssize_t>(npts), static_cast<py::ssize_t>(3)});
        if (!ns.empty()) std::memcpy(num.mutable_data(), ns.data(), sizeof(int32_t) * ns.size());
        float* dst = pts.mutable_data();
        for (const Strand& s : strands)
            if (s.size() >= 6) {
                std::memcpy(dst, s.data(), sizeof(float) * s.size());
                dst += s.size();
            }
        return py::make_tuple(num, pts);
    }, "Load hair strands and return (num_samples, points) pair.");

    m.def("resample_strands_by_arclength", [](py::array ns_in, py::array pts_in, int target) {
        // (N, target, 3) float32 on the host, computed on the current GPU
        // with the kernels of resample_strands_by_arclength_device
        const char* fn = "resample_strands_by_arclength";
        if (ns_in.dtype().kind() != 'i' || ns_in.itemsize() != 4 || ns_in.ndim() != 1)
            throw py::type_error("resample_strands_by_arclength: `num_samples` must be int32 (N,)");
        if (pts_in.dtype().kind() != 'f' || pts_in.itemsize() != 4 || pts_in.ndim() != 2 ||
            pts_in.shape(1) != 3)
            throw py::type_error("resample_strands_by_arclength: `points` must be float32 (P, 3)");
        if (target < 2) throw std::invalid_argument("resample_strands_by_arclength: target < 2");
        py::array_t<int32_t, py::array::c_style | py::array::forcecast> ns(ns_in);
        py::array_t<float, py::array::c_style | py::array::forcecast> pts(pts_in);
        const int64_t N = ns.shape(0), Pn = pts.shape(0);
        std::vector<int64_t> off(static_cast<size_t>(N));
        int64_t sum = 0;
        for (int64_t i = 0; i < N; ++i) {
            if (ns.data()[i] < 2)
                throw std::invalid_argument(
                    "resample_strands_by_arclength: every strand needs at least 2 points");
            off[static_cast<size_t>(i)] = sum;
            sum += ns.data()[i];
        }
        if (sum != Pn)
            throw std::invalid_argument(
                "resample_strands_by_arclength: sum(num_samples) must equal the number of points");
        py::array_t<float> out({static_cast<py::ssize_t>(N), static_cast<py::ssize_t>(target),
                                static_cast<py::ssize_t>(3)});
        if (N == 0) return out;
        const auto dev = at::Device(at::kCUDA, c10::hip::current_device());
        auto f32 = at::TensorOptions().dtype(at::kFloat).device(dev);
        Tensor t_ns = at::from_blob(const_cast<int32_t*>(ns.data()), {N}, at::kInt).to(dev);
        Tensor t_off = at::from_blob(off.data(), {N}, at::kLong).to(dev);
        Tensor t_pts = at::from_blob(const_cast<float*>(pts.data()), {Pn, 3}, at::kFloat).to(dev);
        Tensor lens = at::empty({Pn}, f32);
        Tensor res = at::empty({N, target, 3}, f32);
        {
            DeviceGuard g(dev);
            py::gil_scoped_release nogil;
            check(s2c_resample_strands_by_arclength(I(N, fn), P<int>(t_ns), P<long long>(t_off),
                                                    P<float>(t_pts), target, P<float>(lens),
                                                    P<float>(res), stream_of(t_pts)),
                  fn);
        }
        Tensor host = res.cpu();
        std::memcpy(out.mutable_data(), host.data_ptr<float>(), sizeof(float) * N * target * 3);
        return out;
    }, "Resample strands uniformly by arc length.");

    m.def("resample_strands_by_arclength_device", [](Tensor num_samples, Tensor offsets,
                                                      Tensor points, int target, Tensor lengths,
                                                      Tensor out) {
        const char* fn = "resample_strands_by_arclength";
        need(num_samples, fn, "num_samples", at::kInt, {-1});
        if (offsets.scalar_type() != at::kLong || offsets.dim() != 1 ||
            offsets.size(0) != num_samples.size(0) || !offsets.is_contiguous())
            throw py::type_error("resample_strands_by_arclength: `offsets` must be int64 (N,)");
        need(points, fn, "points", at::kFloat, {-1, 3});
        need(lengths, fn, "lengths", at::kFloat, {points.size(0)});
        need(out, fn, "out", at::kFloat, {num_samples.size(0), target, 3});
        same_device(fn, {&num_samples, &offsets, &points, &lengths, &out});
        DeviceGuard g(points.device());
        check(s2c_resample_strands_by_arclength(I(num_samples.size(0), fn), P<int>(num_samples),
                                                P<long long>(offsets), P<float>(points), target,
                                                P<float>(lengths), P<float>(out),
                                                stream_of(points)),
              fn);
    });

    // Differentiable card coverage (see the C ABI header).
    m.def("card_coverage_forward", [](Tensor pos, Tensor tri, Tensor coverage) {
        const char* fn = "card_coverage_forward";
        need(pos, fn, "pos", at::kFloat, {-1, -1, 4});
        need(tri, fn, "tri", at::kInt, {-1, 3});
        need(coverage, fn, "coverage", at::kFloat, {pos.size(0), -1, -1});
        same_device(fn, {&pos, &tri, &coverage});
        DeviceGuard g(pos.device());
        check(s2c_card_coverage_forward(I(pos.size(0), fn), I(coverage.size(1), fn),
                                        I(coverage.size(2), fn), I(pos.size(1), fn), P<float>(pos),
                                        I(tri.size(0), fn), P<int>(tri), P<float>(coverage),
                                        stream_of(pos)),
              fn);
    });

    m.def("card_coverage_backward", [](Tensor pos, Tensor tri, Tensor d_coverage, Tensor d_pos) {
        const char* fn = "card_coverage_backward";
        need(pos, fn, "pos", at::kFloat, {-1, -1, 4});
        need(tri, fn, "tri", at::kInt, {-1, 3});
        need(d_coverage, fn, "d_coverage", at::kFloat, {pos.size(0), -1, -1});
        need(d_pos, fn, "d_pos", at::kFloat, {pos.size(0), pos.size(1), 4});
        same_device(fn, {&pos, &tri, &d_coverage, &d_pos});
        DeviceGuard g(pos.device());
        check(s2c_card_coverage_backward(I(pos.size(0), fn), I(d_coverage.size(1), fn),
                                         I(d_coverage.size(2), fn), I(pos.size(1), fn),
                                         P<float>(pos), I(tri.size(0), fn), P<int>(tri),
                                         P<float>(d_coverage), P<float>(d_pos), stream_of(pos)),
              fn);
    });

    m.def("triangle_cluster_mask", [](Tensor pos, Tensor tri, Tensor tri2cluster, Tensor bitset) {
        const char* fn = "triangle_cluster_mask";
        need(pos, fn, "pos", at::kFloat, {-1, -1, 4});
        need(tri, fn, "tri", at::kInt, {-1, 3});
        need(tri2cluster, fn, "tri2cluster", at::kInt, {-1});
        need(bitset, fn, "bitset", at::kInt, {-1, -1, -1, -1});
        same_device(fn, {&pos, &tri, &tri2cluster, &bitset});
        if (bitset.size(0) != pos.size(0))
            throw std::runtime_error(
                "triangle_cluster_mask: bitset must have B = pos.shape[0] views");
        if (tri2cluster.size(0) != tri.size(0))
            throw std::runtime_error("triangle_cluster_mask: tri2cluster must have T entries");
        DeviceGuard g(pos.device());
        check(s2c_triangle_cluster_mask(I(pos.size(0), fn), I(bitset.size(1), fn),
                                        I(bitset.size(2), fn), I(pos.size(1), fn), P<float>(pos),
                                        I(tri.size(0), fn), P<int>(tri), P<int>(tri2cluster),
                                        I(bitset.size(3), fn), P<int>(bitset), stream_of(pos)),
              fn);
    });

    // Grid-mesh large-steps system (no reference binding: replaces
    // largesteps.parameterize.to_differential / from_differential).
    auto grid_layout = [](const char* fn, const Tensor& rows, const Tensor& cols,
                          const Tensor& voff, const Tensor& a, const Tensor& b) {
        need(rows, fn, "rows", at::kInt, {-1});
        need(cols, fn, "cols", at::kInt, {rows.size(0)});
        need(voff, fn, "voff", at::kInt, {rows.size(0)});
        need(a, fn, "input", at::kFloat, {-1, 3});
        need(b, fn, "output", at::kFloat, {a.size(0), 3});
        same_device(fn, {&rows, &cols, &voff, &a, &b});
        return I(rows.size(0), fn);
    };
    m.def("grid_system_apply", [grid_layout](Tensor rows, Tensor cols, Tensor voff, float lam,
                                             float w_weight, Tensor x, Tensor u) {
        const char* fn = "grid_system_apply";
        const int G = grid_layout(fn, rows, cols, voff, x, u);
        DeviceGuard g(x.device());
        check(s2c_grid_system_apply(G, P<int>(rows), P<int>(cols), P<int>(voff), lam, w_weight,
                                    P<float>(x), P<float>(u), stream_of(x)),
              fn);
    });
    m.def("grid_laplacian", [grid_layout](Tensor rows, Tensor cols, Tensor voff, float w_weight,
                                          Tensor x, Tensor y) {
        const char* fn = "grid_laplacian";
        const int G = grid_layout(fn, rows, cols, voff, x, y);
        DeviceGuard g(x.device());
        check(s2c_grid_laplacian(G, P<int>(rows), P<int>(cols), P<int>(voff), w_weight,
                                 P<float>(x), P<float>(y), stream_of(x)),
              fn);
    });
    m.def("grid_quad_diagonals", [](Tensor rows, Tensor cols, Tensor voff, Tensor qoff,
                                    Tensor pos, Tensor D, bool backward) {
        const char* fn = "grid_quad_diagonals";
        need(rows, fn, "rows", at::kInt, {-1});
        need(cols, fn, "cols", at::kInt, {rows.size(0)});
        need(voff, fn, "voff", at::kInt, {rows.size(0)});
        need(qoff, fn, "qoff", at::kInt, {rows.size(0)});
        need(pos, fn, "pos", at::kFloat, {-1, 3});
        need(D, fn, "D", at::kFloat, {-1, 3});
        same_device(fn, {&rows, &cols, &voff, &qoff, &pos, &D});
        DeviceGuard g(pos.device());
        const int G = I(rows.size(0), fn);
        if (backward)  // D holds dD, pos receives dpos
            check(s2c_grid_quad_diagonals_backward(G, P<int>(rows), P<int>(cols), P<int>(voff),
                                                   P<int>(qoff), P<float>(D), P<float>(pos),
                                                   stream_of(pos)),
                  fn);
        else
            check(s2c_grid_quad_diagonals(G, P<int>(rows), P<int>(cols), P<int>(voff),
                                          P<int>(qoff), P<float>(pos), P<float>(D),
                                          stream_of(pos)),
                  fn);
    });
    m.def("grid_system_factor", [](Tensor rows, Tensor cols, Tensor fac_off, float lam,
                                   float w_weight, Tensor factor) {
        const char* fn = "grid_system_factor";
        need(rows, fn, "rows", at::kInt, {-1});
        need(cols, fn, "cols", at::kInt, {rows.size(0)});
        need(fac_off, fn, "fac_off", at::kInt, {rows.size(0)});
        need(factor, fn, "factor", at::kFloat, {-1});
        same_device(fn, {&rows, &cols, &fac_off, &factor});
        DeviceGuard g(factor.device());
        check(s2c_grid_system_factor(I(rows.size(0), fn), P<int>(rows), P<int>(cols),
                                     P<int>(fac_off), lam, w_weight, P<float>(factor),
                                     stream_of(factor)),
              fn);
    });
    m.def("grid_system_solve", [grid_layout](Tensor rows, Tensor cols, Tensor voff,
                                             Tensor fac_off, Tensor factor, int max_verts,
                                             int max_rows, float lam, Tensor u, Tensor x) {
        const char* fn = "grid_system_solve";
        const int G = grid_layout(fn, rows, cols, voff, u, x);
        need(fac_off, fn, "fac_off", at::kInt, {rows.size(0)});
        need(factor, fn, "factor", at::kFloat, {-1});
        same_device(fn, {&rows, &fac_off, &factor});
        DeviceGuard g(u.device());
        check(s2c_grid_system_solve(G, P<int>(rows), P<int>(cols), P<int>(voff), P<int>(fac_off),
                                    P<float>(factor), max_verts, max_rows, lam, P<float>(u),
                                    P<float>(x), stream_of(u)),
              fn);
    });

    // Card mesh inputs (no reference binding: the scripts build these on
    // the host, 2_texture.py:17-96, texture_utils.py:525-655,
    // 3_fitting.py:17-60).
    m.def("strands_to_strips", [](Tensor counts, Tensor offsets, Tensor points, int res_w,
                                  double width, std::vector<float> rot_deg, Tensor out) {
        const char* fn = "strands_to_strips";
        need(counts, fn, "counts", at::kInt, {-1});
        need(offsets, fn, "offsets", at::kLong, {counts.size(0)});
        need(points, fn, "points", at::kFloat, {-1, 3});
        need(out, fn, "out", at::kFloat, {-1, 3});
        same_device(fn, {&counts, &offsets, &points, &out});
        if (out.size(0) != points.size(0) * res_w * static_cast<int64_t>(rot_deg.size()))
            throw std::runtime_error("strands_to_strips: out must hold res_w * rotations * points rows");
        DeviceGuard g(points.device());
        check(s2c_strands_to_strips(I(counts.size(0), fn), P<int>(counts), P<long long>(offsets),
                                    P<float>(points), res_w, width, I(rot_deg.size(), fn),
                                    rot_deg.data(), P<float>(out), stream_of(points)),
              fn);
    });
    m.def("grid_mesh_topology", [](Tensor rows, Tensor cols, Tensor voff, Tensor qoff, Tensor hoff,
                                   Tensor woff, std::optional<Tensor> grid2cluster,
                                   std::optional<Tensor> rects, double margin_top, Tensor faces,
                                   Tensor quads, Tensor h_edges, Tensor w_edges,
                                   std::optional<Tensor> face2cluster, std::optional<Tensor> uvs) {
        const char* fn = "grid_mesh_topology";
        need(rows, fn, "rows", at::kInt, {-1});
        const int64_t G = rows.size(0);
        for (auto* t : {&cols, &voff, &qoff, &hoff, &woff}) need(*t, fn, "layout", at::kInt, {G});
        need(faces, fn, "faces", at::kInt, {-1, 3});
        need(quads, fn, "quads", at::kInt, {-1, 4});
        need(h_edges, fn, "h_edges", at::kInt, {-1, 2});
        need(w_edges, fn, "w_edges", at::kInt, {-1, 2});
        if (faces.size(0) != 2 * quads.size(0))
            throw std::runtime_error("grid_mesh_topology: faces must have 2 rows per quad");
        same_device(fn, {&rows, &cols, &voff, &qoff, &hoff, &woff, &faces, &quads, &h_edges, &w_edges});
        if (grid2cluster.has_value() != face2cluster.has_value() || rects.has_value() != uvs.has_value())
            throw std::runtime_error("grid_mesh_topology: grid2cluster/face2cluster and rects/uvs go in pairs");
        if (grid2cluster) {
            need(*grid2cluster, fn, "grid2cluster", at::kLong, {G});
            need(*face2cluster, fn, "face2cluster", at::kLong, {faces.size(0)});
            same_device(fn, {&rows, &*grid2cluster, &*face2cluster});
        }
        if (rects) {
            need(*rects, fn, "rects", at::kDouble, {G, 4});
            need(*uvs, fn, "uvs", at::kDouble, {-1, 2});
            same_device(fn, {&rows, &*rects, &*uvs});
        }
        DeviceGuard g(rows.device());
        check(s2c_grid_mesh_topology(
                  I(G, fn), P<int>(rows), P<int>(cols), P<int>(voff), P<int>(qoff), P<int>(hoff),
                  P<int>(woff), grid2cluster ? P<long long>(*grid2cluster) : nullptr,
                  rects ? P<double>(*rects) : nullptr, margin_top, P<int>(faces), P<int>(quads),
                  P<int>(h_edges), P<int>(w_edges),
                  face2cluster ? P<long long>(*face2cluster) : nullptr,
                  uvs ? P<double>(*uvs) : nullptr, stream_of(rows)),
              fn);
    });
    m.def("strand_edges", [](Tensor counts, Tensor offsets, Tensor cluster_ids, Tensor edges,
                             Tensor edge2cluster) {
        const char* fn = "strand_edges";
        need(counts, fn, "counts", at::kInt, {-1});
        need(offsets, fn, "offsets", at::kLong, {counts.size(0)});
        need(cluster_ids, fn, "cluster_ids", at::kInt, {counts.size(0)});
        need(edges, fn, "edges", at::kInt, {-1, 2});
        need(edge2cluster, fn, "edge2cluster", at::kInt, {edges.size(0)});
        same_device(fn, {&counts, &offsets, &cluster_ids, &edges, &edge2cluster});
        DeviceGuard g(counts.device());
        check(s2c_strand_edges(I(counts.size(0), fn), P<int>(counts), P<long long>(offsets),
                               P<int>(cluster_ids), P<int>(edges), P<int>(edge2cluster),
                               stream_of(counts)),
              fn);
    });

    // Software-MSAA texture patches (no reference binding: render_texture,
    // texture_utils.py:106-288, loops over rasterize on the host).
    m.def("msaa_jitter", [](Tensor pos, Tensor dx, Tensor dy, Tensor out) {
        const char* fn = "msaa_jitter";
        need(pos, fn, "pos", at::kFloat, {-1, 3});
        need(dx, fn, "dx", at::kFloat, {-1});
        need(dy, fn, "dy", at::kFloat, {dx.size(0)});
        need(out, fn, "out", at::kFloat, {dx.size(0), pos.size(0), 4});
        same_device(fn, {&pos, &dx, &dy, &out});
        DeviceGuard g(pos.device());
        check(s2c_msaa_jitter(I(pos.size(0), fn), P<float>(pos), I(dx.size(0), fn), P<float>(dx),
                              P<float>(dy), P<float>(out), stream_of(pos)),
              fn);
    });
    m.def("msaa_resolve", [](Tensor rast, Tensor attr, Tensor tri, Tensor weights, Tensor rgb,
                             Tensor alpha) {
        const char* fn = "msaa_resolve";
        need(rast, fn, "rast", at::kFloat, {-1, -1, -1, 4});
        need(attr, fn, "attr", at::kFloat, {-1, 3});
        need(tri, fn, "tri", at::kInt, {-1, 3});
        need(weights, fn, "weights", at::kFloat, {rast.size(0)});
        need(rgb, fn, "rgb", at::kFloat, {rast.size(1), rast.size(2), 3});
        need(alpha, fn, "alpha", at::kFloat, {rast.size(1), rast.size(2)});
        same_device(fn, {&rast, &attr, &tri, &weights, &rgb, &alpha});
        DeviceGuard g(rast.device());
        check(s2c_msaa_resolve(I(rast.size(0), fn), I(rast.size(1), fn), I(rast.size(2), fn),
                               P<float>(rast), P<float>(attr), P<int>(tri), P<float>(weights),
                               P<float>(rgb), P<float>(alpha), stream_of(rast)),
              fn);
    });
    m.def("blur3x3_reflect", [](std::vector<float> kernel3, Tensor in, Tensor tmp, Tensor out) {
        const char* fn = "blur3x3_reflect";
        need(in, fn, "in", at::kFloat, {-1, -1});
        need(tmp, fn, "tmp", at::kFloat, {in.size(0), in.size(1)});
        need(out, fn, "out", at::kFloat, {in.size(0), in.size(1)});
        same_device(fn, {&in, &tmp, &out});
        if (kernel3.size() != 3) throw std::runtime_error("blur3x3_reflect: kernel3 needs 3 taps");
        DeviceGuard g(in.device());
        check(s2c_blur3x3_reflect(I(in.size(0), fn), I(in.size(1), fn), kernel3.data(),
                                  P<float>(in), P<float>(tmp), P<float>(out), stream_of(in)),
              fn);
    });
    m.def("atlas_paste", [](Tensor rgb, Tensor alpha, int crop_x0, int crop_y0, int atlas_x0,
                            int atlas_y0, int atlas_x1, int atlas_y1, Tensor atlas_rgb,
                            Tensor atlas_alpha) {
        const char* fn = "atlas_paste";
        need(rgb, fn, "rgb", at::kFloat, {-1, -1, 3});
        need(alpha, fn, "alpha", at::kFloat, {rgb.size(0), rgb.size(1)});
        need(atlas_rgb, fn, "atlas_rgb", at::kByte, {-1, -1, 3});
        need(atlas_alpha, fn, "atlas_alpha", at::kByte, {atlas_rgb.size(0), atlas_rgb.size(1)});
        same_device(fn, {&rgb, &alpha, &atlas_rgb, &atlas_alpha});
        if (atlas_rgb.size(0) != atlas_rgb.size(1))
            throw std::runtime_error("atlas_paste: the atlas must be square");
        DeviceGuard g(rgb.device());
        check(s2c_atlas_paste(I(rgb.size(0), fn), I(rgb.size(1), fn), P<float>(rgb),
                              P<float>(alpha), crop_x0, crop_y0, atlas_x0, atlas_y0, atlas_x1,
                              atlas_y1, I(atlas_rgb.size(0), fn), P<unsigned char>(atlas_rgb),
                              P<unsigned char>(atlas_alpha), stream_of(rgb)),
              fn);
    });
    m.def("msaa_patch", [](Tensor jittered, Tensor tri, Tensor attr, Tensor weights, int H, int W,
                           Tensor keys, Tensor rgb, Tensor alpha) {
        const char* fn = "msaa_patch";
        need(jittered, fn, "jittered", at::kFloat, {-1, -1, 4});
        need(tri, fn, "tri", at::kInt, {-1, 3});
        need(attr, fn, "attr", at::kFloat, {jittered.size(1), 3});
        need(weights, fn, "weights", at::kFloat, {jittered.size(0)});
        need(keys, fn, "keys", at::kLong, {jittered.size(0), H, W});
        need(rgb, fn, "rgb", at::kFloat, {H, W, 3});
        need(alpha, fn, "alpha", at::kFloat, {H, W});
        same_device(fn, {&jittered, &tri, &attr, &weights, &keys, &rgb, &alpha});
        DeviceGuard g(jittered.device());
        check(s2c_msaa_patch(I(jittered.size(1), fn), I(tri.size(0), fn), P<float>(jittered),
                             P<int>(tri), P<float>(attr), I(jittered.size(0), fn), H, W,
                             P<float>(weights), P<long long>(keys), P<float>(rgb),
                             P<float>(alpha), stream_of(jittered)),
              fn);
    });
    m.def("patch_masks", [](Tensor rows, Tensor cols, Tensor voff, int max_edges, Tensor pos,
                            Tensor keys, Tensor out) {
        const char* fn = "patch_masks";
        need(rows, fn, "rows", at::kInt, {-1});
        need(cols, fn, "cols", at::kInt, {rows.size(0)});
        need(voff, fn, "voff", at::kLong, {rows.size(0)});
        need(pos, fn, "pos", at::kFloat, {-1, 4});
        need(out, fn, "out", at::kByte, {rows.size(0), -1, -1});
        need(keys, fn, "keys", at::kLong, {rows.size(0), out.size(1), out.size(2)});
        same_device(fn, {&rows, &cols, &voff, &pos, &keys, &out});
        DeviceGuard g(pos.device());
        check(s2c_patch_masks(I(rows.size(0), fn), I(out.size(1), fn), I(out.size(2), fn),
                              P<int>(rows), P<int>(cols), P<long long>(voff), max_edges,
                              P<float>(pos), P<long long>(keys), P<unsigned char>(out),
                              stream_of(pos)),
              fn);
    });
}
