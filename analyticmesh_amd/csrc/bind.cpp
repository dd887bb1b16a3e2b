// pybind11 module `_am`: a thin object wrapper over the C-ABI of include/am.h.
// Buffers cross as integer addresses (host or device), so no torch types
// appear in the native signatures.
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/am.h"

namespace py = pybind11;

namespace {

void check(am_ctx* c, am_status s) {
  if (s != AM_OK) throw std::runtime_error("am error " + std::to_string(s) + ": " + am_last_error(c));
}

struct Engine {
  am_ctx* c = nullptr;
  int vm = 20;
  int n_linear = 0;

  Engine(const std::vector<int>& nodes, py::array_t<int, py::array::c_style | py::array::forcecast> arc,
         int num_extra, int precision) {
    if (arc.ndim() != 2) throw std::runtime_error("arc_table must be 2-D");
    const am_status s = am_create(&c, (int)nodes.size(), nodes.data(), arc.data(), (int)arc.shape(0),
                                  (int)arc.shape(1), num_extra, precision);
    if (s != AM_OK) throw std::runtime_error("am error " + std::to_string(s) + ": " + am_last_error(nullptr));
    n_linear = (int)nodes.size() - 1;
  }
  ~Engine() {
    if (c) am_destroy(c);
  }

  void set_option(const std::string& key, long long value) {
    check(c, am_set_option(c, key.c_str(), value));
    if (key == "vm") vm = (int)value;
  }

  void run(const std::vector<uintptr_t>& weights, const std::vector<uintptr_t>& biases,
           const std::vector<std::pair<int, int>>& tm_shapes, const std::vector<uintptr_t>& tms, int dtype,
           uintptr_t points, uintptr_t states, int n_seeds, uintptr_t extra_w, uintptr_t extra_b, double iso,
           bool flip, uintptr_t stream) {
    std::vector<const void*> w(weights.size()), b(biases.size()), t(tms.size());
    for (size_t i = 0; i < w.size(); ++i) w[i] = (const void*)weights[i];
    for (size_t i = 0; i < b.size(); ++i) b[i] = (const void*)biases[i];
    for (size_t i = 0; i < t.size(); ++i) t[i] = (const void*)tms[i];
    std::vector<int> shp;
    for (auto& s : tm_shapes) {
      shp.push_back(s.first);
      shp.push_back(s.second);
    }
    am_status s;
    {
      py::gil_scoped_release nogil;
      s = am_run(c, w.data(), b.data(), (int)t.size(), shp.data(), t.data(), dtype, (const void*)points,
                 (const uint8_t*)states, n_seeds, (const void*)extra_w, (const void*)extra_b, iso, flip ? 1 : 0,
                 (void*)stream);
    }
    check(c, s);
  }

  void combine(double scale, const std::vector<double>& center) {
    if (center.size() != 3) throw std::runtime_error("center must have 3 entries");
    am_status s;
    {
      py::gil_scoped_release nogil;
      s = am_combine(c, scale, center.data());
    }
    check(c, s);
  }

  void export_ply(const std::string& path, bool is_poly, bool is_f32) {
    check(c, am_export_ply(c, path.c_str(), is_poly ? 1 : 0, is_f32 ? 1 : 0));
  }

  py::tuple counts() {
    long long r, f, v, n;
    check(c, am_get_counts(c, &r, &f, &v, &n));
    return py::make_tuple(r, f, v, n);
  }

  py::tuple mesh() {
    long long r, f, v, n;
    check(c, am_get_counts(c, &r, &f, &v, &n));
    py::array_t<double> V({(py::ssize_t)v, (py::ssize_t)3});
    py::array_t<int> off((py::ssize_t)(f + 1));
    py::array_t<int> idx((py::ssize_t)n);
    check(c, am_get_mesh(c, V.mutable_data(), off.mutable_data(), idx.mutable_data()));
    return py::make_tuple(V, off, idx);
  }

  py::dict regions(int Wd) {
    long long r, f, v, n;
    check(c, am_get_counts(c, &r, &f, &v, &n));
    py::array_t<uint32_t> masks({(py::ssize_t)r, (py::ssize_t)Wd});
    py::array_t<int> info({(py::ssize_t)r, (py::ssize_t)8});
    py::array_t<int> tabs({(py::ssize_t)r, (py::ssize_t)(vm + 1)});
    py::array_t<double> verts({(py::ssize_t)r, (py::ssize_t)vm, (py::ssize_t)3});
    py::array_t<int> vpairs({(py::ssize_t)r, (py::ssize_t)vm, (py::ssize_t)2});
    py::array_t<double> points({(py::ssize_t)r, (py::ssize_t)3});
    check(c, am_get_regions(c, masks.mutable_data(), info.mutable_data(), tabs.mutable_data(), verts.mutable_data(),
                            vpairs.mutable_data(), points.mutable_data()));
    py::dict d;
    d["masks"] = masks;
    d["info"] = info;
    d["tabs"] = tabs;
    d["verts"] = verts;
    d["vpairs"] = vpairs;
    d["points"] = points;
    return d;
  }

  void adjoint(uintptr_t gradv, int dtype, const std::vector<uintptr_t>& gw, const std::vector<uintptr_t>& gb) {
    // am_adjoint writes one array per linear layer into each list
    if (gw.size() != (size_t)n_linear || gb.size() != (size_t)n_linear)
      throw std::runtime_error("adjoint needs one weight and one bias gradient buffer per linear layer (" +
                               std::to_string(n_linear) + ")");
    std::vector<void*> w(gw.size()), b(gb.size());
    for (size_t i = 0; i < w.size(); ++i) w[i] = (void*)gw[i];
    for (size_t i = 0; i < b.size(); ++i) b[i] = (void*)gb[i];
    am_status s;
    {
      py::gil_scoped_release nogil;
      s = am_adjoint(c, (const void*)gradv, dtype, w.data(), b.data());
    }
    check(c, s);
  }

  py::tuple planes(py::array_t<uint32_t, py::array::c_style | py::array::forcecast> masks, int C) {
    const int n = (int)masks.shape(0);
    py::array_t<float> P({(py::ssize_t)n, (py::ssize_t)4, (py::ssize_t)C});
    py::array_t<float> E({(py::ssize_t)n, (py::ssize_t)4});
    py::array_t<float> S({(py::ssize_t)n, (py::ssize_t)(C + 1)});
    check(c, am_planes(c, masks.data(), n, P.mutable_data(), E.mutable_data(), S.mutable_data()));
    return py::make_tuple(P, E, S);
  }

  py::dict fingerprint() {
    unsigned long long f[5];
    check(c, am_topology_fingerprint(c, f));
    long long r, nf, nv, ni;
    check(c, am_get_counts(c, &r, &nf, &nv, &ni));
    char buf[3][17];
    for (int i = 0; i < 3; ++i) std::snprintf(buf[i], sizeof(buf[i]), "%016llx", f[i]);
    py::dict d;
    d["regions"] = r;
    d["faces"] = (long long)f[3];
    d["vertex_uses"] = (long long)f[4];
    d["visited"] = std::string(buf[0]);
    d["faces_hash"] = std::string(buf[1]);
    d["vertex_keys"] = std::string(buf[2]);
    return d;
  }

  uintptr_t ctx_address() const { return (uintptr_t)c; }

  py::array_t<double> stats() {
    py::array_t<double> s((py::ssize_t)AM_STAT_COUNT);
    check(c, am_get_stats(c, s.mutable_data(), AM_STAT_COUNT));
    return s;
  }
};

struct Batch {
  am_batch* b = nullptr;
  int vm = 20;
  std::vector<int> Wn;

  Batch(const std::vector<std::vector<int>>& nets, int precision) {
    std::vector<int> nn;
    std::vector<const int*> ptrs;
    for (auto& n : nets) {
      nn.push_back((int)n.size());
      ptrs.push_back(n.data());
      int L = 0;
      for (size_t k = 1; k + 1 < n.size(); ++k) L += n[k];
      Wn.push_back((L + 31) / 32);
    }
    const am_status s = am_batch_create(&b, (int)nets.size(), nn.data(), ptrs.data(), precision);
    if (s != AM_OK) throw std::runtime_error("am error " + std::to_string(s) + ": " + am_last_error(nullptr));
  }
  ~Batch() {
    if (b) am_batch_destroy(b);
  }
  void chk(am_status s) {
    if (s != AM_OK) throw std::runtime_error("am error " + std::to_string(s) + ": " + am_batch_last_error(b));
  }
  void set_option(const std::string& key, long long value) {
    chk(am_batch_set_option(b, key.c_str(), value));
    if (key == "vm") vm = (int)value;
  }
  void run(const std::vector<uintptr_t>& weights, const std::vector<uintptr_t>& biases, int dtype,
           const std::vector<uintptr_t>& points, const std::vector<uintptr_t>& states, const std::vector<int>& n_seeds,
           double iso, bool flip, uintptr_t stream) {
    std::vector<const void*> w(weights.size()), bb(biases.size()), p(points.size());
    std::vector<const uint8_t*> st(states.size());
    for (size_t i = 0; i < w.size(); ++i) w[i] = (const void*)weights[i];
    for (size_t i = 0; i < bb.size(); ++i) bb[i] = (const void*)biases[i];
    for (size_t i = 0; i < p.size(); ++i) p[i] = (const void*)points[i];
    for (size_t i = 0; i < st.size(); ++i) st[i] = (const uint8_t*)states[i];
    if (p.size() != Wn.size() || st.size() != Wn.size() || n_seeds.size() != Wn.size())
      throw std::runtime_error("one seed set per network");
    am_status s;
    {
      py::gil_scoped_release nogil;
      s = am_batch_run(b, w.data(), bb.data(), dtype, p.data(), st.data(), n_seeds.data(), iso, flip ? 1 : 0,
                       (void*)stream);
    }
    chk(s);
  }
  py::dict fingerprint(int net) {
    unsigned long long f[6];
    chk(am_batch_fingerprint(b, net, f));
    char buf[3][17];
    for (int i = 0; i < 3; ++i) std::snprintf(buf[i], sizeof(buf[i]), "%016llx", f[i]);
    py::dict d;
    d["regions"] = (long long)f[5];
    d["faces"] = (long long)f[3];
    d["vertex_uses"] = (long long)f[4];
    d["visited"] = std::string(buf[0]);
    d["faces_hash"] = std::string(buf[1]);
    d["vertex_keys"] = std::string(buf[2]);
    return d;
  }
  py::dict regions(int net) {
    if (net < 0 || net >= (int)Wn.size()) throw std::runtime_error("no such network");
    long long r = 0;
    chk(am_batch_get_regions(b, net, &r, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr));
    py::array_t<uint32_t> masks({(py::ssize_t)r, (py::ssize_t)Wn[net]});
    py::array_t<int> info({(py::ssize_t)r, (py::ssize_t)8});
    py::array_t<int> tabs({(py::ssize_t)r, (py::ssize_t)(vm + 1)});
    py::array_t<double> verts({(py::ssize_t)r, (py::ssize_t)vm, (py::ssize_t)3});
    py::array_t<int> vpairs({(py::ssize_t)r, (py::ssize_t)vm, (py::ssize_t)2});
    py::array_t<double> points({(py::ssize_t)r, (py::ssize_t)3});
    if (r)
      chk(am_batch_get_regions(b, net, &r, masks.mutable_data(), info.mutable_data(), tabs.mutable_data(),
                               verts.mutable_data(), vpairs.mutable_data(), points.mutable_data()));
    py::dict d;
    d["masks"] = masks;
    d["info"] = info;
    d["tabs"] = tabs;
    d["verts"] = verts;
    d["vpairs"] = vpairs;
    d["points"] = points;
    return d;
  }
  void combine(double scale, const std::vector<double>& center) {
    if (center.size() != 3) throw std::runtime_error("center must have 3 entries");
    am_status s;
    {
      py::gil_scoped_release nogil;
      s = am_batch_combine(b, scale, center.data());
    }
    chk(s);
  }
  py::tuple counts() {
    long long r, f, v, n;
    chk(am_batch_get_counts(b, &r, &f, &v, &n));
    return py::make_tuple(r, f, v, n);
  }
  py::array_t<double> stats() {
    py::array_t<double> s((py::ssize_t)AM_STAT_COUNT);
    chk(am_batch_get_stats(b, s.mutable_data(), AM_STAT_COUNT));
    return s;
  }
};

}  // namespace

PYBIND11_MODULE(_am, m) {
  m.doc() = "MI355X analytic-marching engine (C-ABI wrapper)";
  m.attr("PREC_FP32") = AM_PREC_FP32;
  m.attr("PREC_SPLIT_BF16X3") = AM_PREC_SPLIT_BF16X3;
  m.attr("PREC_FP64") = AM_PREC_FP64;
  m.attr("PREC_FP64_RAY") = AM_PREC_FP64_RAY;
  m.attr("DTYPE_F32") = AM_DTYPE_F32;
  m.attr("DTYPE_F64") = AM_DTYPE_F64;
  py::class_<Engine>(m, "Engine")
      .def(py::init<const std::vector<int>&, py::array_t<int, py::array::c_style | py::array::forcecast>, int, int>(),
           py::arg("nodes"), py::arg("arc_table"), py::arg("num_extra_constraints"), py::arg("precision"))
      .def("set_option", &Engine::set_option)
      .def("run", &Engine::run, py::arg("weights"), py::arg("biases"), py::arg("tm_shapes"), py::arg("tms"),
           py::arg("dtype"), py::arg("points"), py::arg("states"), py::arg("n_seeds"), py::arg("extra_w"),
           py::arg("extra_b"), py::arg("iso"), py::arg("flip"), py::arg("stream"))
      .def("combine", &Engine::combine)
      .def("export_ply", &Engine::export_ply)
      .def("counts", &Engine::counts)
      .def("mesh", &Engine::mesh)
      .def("regions", &Engine::regions)
      .def("adjoint", &Engine::adjoint)
      .def("planes", &Engine::planes)
      .def("fingerprint", &Engine::fingerprint)
      .def("ctx_address", &Engine::ctx_address)
      .def("stats", &Engine::stats);
  py::class_<Batch>(m, "Batch")
      .def(py::init<const std::vector<std::vector<int>>&, int>(), py::arg("nets"), py::arg("precision"))
      .def("set_option", &Batch::set_option)
      .def("run", &Batch::run)
      .def("fingerprint", &Batch::fingerprint)
      .def("regions", &Batch::regions)
      .def("combine", &Batch::combine)
      .def("counts", &Batch::counts)
      .def("stats", &Batch::stats);
}
