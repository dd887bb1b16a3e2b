// Polygon-mesh file tools (pybind11 module `_polylib`): the post-processing
// surface of the reference's libpolytools (backend/libpolytools/src/polylib.cpp:
// 53-571), written for this engine's meshes.
//
// PolyMesh holds float vertices, polygon faces and optional per-face RGB colours:
//  * load/save: binary little-endian PLY (float or double x/y/z on load, float
//    on save; optional uchar red/green/blue per face) and ASCII OFF (colours as
//    0..1 floats after the indices on load, written as integers on save);
//  * poly2tri: fan triangulation (0, j+1, j+2) dropping triangles with a
//    repeated index, colours following their faces;
//  * is_polymesh / num_polyfaces / num_trifaces (fan count, degenerate
//    triangles included);
//  * weld(tol): merges vertices whose coordinates agree within tol (per axis),
//    e.g. across the shared faces of voxel-mode pieces; not in the reference.
// Errors are std::runtime_error (Python RuntimeError), as in the reference.
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <algorithm>
#include <array>
#include <cctype>
#include <cmath>
#include <cstdint>
#include <fstream>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

namespace py = pybind11;

namespace {

std::string lower(std::string s) {
  for (char& ch : s) ch = (char)std::tolower((unsigned char)ch);
  return s;
}

std::string extension_of(const std::string& path) {
  const size_t dot = path.find_last_of('.');
  return dot == std::string::npos ? std::string() : lower(path.substr(dot + 1));
}

std::vector<std::string> words(const std::string& line) {
  std::istringstream is(line);
  std::vector<std::string> out;
  std::string w;
  while (is >> w) out.push_back(w);
  return out;
}

class PolyMesh {
 public:
  PolyMesh() = default;
  explicit PolyMesh(const std::string& path) { load(path); }
  PolyMesh(const std::vector<std::vector<float>>& v, const std::vector<std::vector<int>>& f,
           const std::vector<std::vector<int>>& c) {
    set_vertices(v);
    set_faces(f);
    set_colors(c);
  }

  void clear() {
    v_.clear();
    f_.clear();
    c_.clear();
  }

  void load(const std::string& path) {
    const std::string ext = extension_of(path);
    if (ext == "ply") load_ply(path);
    else if (ext == "off") load_off(path);
    else throw std::runtime_error("Unsupported file format: " + path);
  }

  void save(const std::string& path) const {
    const std::string ext = extension_of(path);
    if (ext == "ply") save_ply(path);
    else if (ext == "off") save_off(path);
    else throw std::runtime_error("Unsupported file format: " + path);
  }

  std::vector<std::vector<float>> vertices() const {
    std::vector<std::vector<float>> out;
    out.reserve(v_.size());
    for (const auto& p : v_) out.push_back({p[0], p[1], p[2]});
    return out;
  }
  std::vector<std::vector<int>> faces() const { return f_; }
  std::vector<std::vector<int>> colors() const {
    std::vector<std::vector<int>> out;
    out.reserve(c_.size());
    for (const auto& c : c_) out.push_back({c[0], c[1], c[2]});
    return out;
  }

  void set_vertices(const std::vector<std::vector<float>>& v) {
    for (const auto& p : v)
      if (p.size() != 3) throw std::runtime_error("Vertices should be 3-dimensional.");
    v_.clear();
    v_.reserve(v.size());
    for (const auto& p : v) v_.push_back({p[0], p[1], p[2]});
  }
  void set_faces(const std::vector<std::vector<int>>& f) {
    for (const auto& x : f)
      if (x.size() < 3) throw std::runtime_error("Faces should have at least 3 vertices.");
    f_ = f;
  }
  void set_colors(const std::vector<std::vector<int>>& c) {
    for (const auto& x : c)
      if (x.size() != 3 || x[0] < 0 || x[0] > 255 || x[1] < 0 || x[1] > 255 || x[2] < 0 || x[2] > 255)
        throw std::runtime_error("RGB colors should be in integer range: [0, 255].");
    c_.clear();
    for (const auto& x : c) c_.push_back({(uint8_t)x[0], (uint8_t)x[1], (uint8_t)x[2]});
  }

  void poly2tri() {
    std::vector<std::vector<int>> tf;
    std::vector<std::array<uint8_t, 3>> tc;
    const bool colored = !c_.empty();
    for (size_t i = 0; i < f_.size(); ++i) {
      const auto& f = f_[i];
      for (size_t j = 1; j + 1 < f.size(); ++j) {
        const int a = f[0], b = f[j], c = f[j + 1];
        if (a == b || b == c || c == a) continue;   // degenerate triangle
        tf.push_back({a, b, c});
        if (colored) tc.push_back(c_[i]);
      }
    }
    f_.swap(tf);
    c_.swap(tc);
  }

  bool is_polymesh() const {
    if (f_.empty()) return true;
    for (const auto& f : f_)
      if (f.size() > 3) return true;
    return false;
  }
  long long num_polyfaces() const { return (long long)f_.size(); }
  long long num_trifaces() const {
    if (!is_polymesh()) return (long long)f_.size();
    long long n = 0;
    for (const auto& f : f_) n += (long long)f.size() - 2;
    return n;
  }

  // Merge vertices within tol per axis (first occurrence kept, faces re-indexed;
  // faces that collapse below three distinct vertices are dropped).  Returns the
  // number of vertices removed.
  long long weld(double tol) {
    if (!(tol > 0)) throw std::runtime_error("weld tolerance must be positive");
    std::unordered_map<long long, std::vector<int>> grid;   // cell hash -> kept vertex ids
    auto cell = [&](float x) { return (long long)std::floor((double)x / tol); };
    auto key = [](long long a, long long b, long long c) {
      return (long long)((uint64_t)a * 73856093ULL ^ (uint64_t)b * 19349663ULL ^ (uint64_t)c * 83492791ULL);
    };
    std::vector<int> remap(v_.size());
    std::vector<std::array<float, 3>> kept;
    for (size_t i = 0; i < v_.size(); ++i) {
      const auto& p = v_[i];
      const long long cx = cell(p[0]), cy = cell(p[1]), cz = cell(p[2]);
      int found = -1;
      for (long long dx = -1; dx <= 1 && found < 0; ++dx)
        for (long long dy = -1; dy <= 1 && found < 0; ++dy)
          for (long long dz = -1; dz <= 1 && found < 0; ++dz) {
            auto it = grid.find(key(cx + dx, cy + dy, cz + dz));
            if (it == grid.end()) continue;
            for (int k : it->second) {
              const auto& q = kept[k];
              if (std::fabs(q[0] - p[0]) <= tol && std::fabs(q[1] - p[1]) <= tol && std::fabs(q[2] - p[2]) <= tol) {
                found = k;
                break;
              }
            }
          }
      if (found < 0) {
        found = (int)kept.size();
        kept.push_back(p);
        grid[key(cx, cy, cz)].push_back(found);
      }
      remap[i] = found;
    }
    const long long removed = (long long)v_.size() - (long long)kept.size();
    std::vector<std::vector<int>> nf;
    std::vector<std::array<uint8_t, 3>> nc;
    for (size_t i = 0; i < f_.size(); ++i) {
      std::vector<int> f;
      for (int x : f_[i]) {
        if (x < 0 || x >= (int)remap.size()) throw std::runtime_error("face index out of range");
        const int y = remap[x];
        if (f.empty() || f.back() != y) f.push_back(y);
      }
      while (f.size() > 1 && f.front() == f.back()) f.pop_back();
      if (f.size() < 3) continue;
      nf.push_back(std::move(f));
      if (!c_.empty()) nc.push_back(c_[i]);
    }
    v_.swap(kept);
    f_.swap(nf);
    c_.swap(nc);
    return removed;
  }

 private:
  std::vector<std::array<float, 3>> v_;
  std::vector<std::vector<int>> f_;
  std::vector<std::array<uint8_t, 3>> c_;

  void load_ply(const std::string& path) {
    std::ifstream in(path, std::ios::binary);
    if (!in) throw std::runtime_error("cannot open " + path);
    std::string line;
    std::getline(in, line);
    if (lower(line) != "ply") throw std::runtime_error("It should be a ply file!");
    bool le = false, have_vertex = false, have_face = false, dbl = false;
    long long nv = 0, nf = 0;
    int face_colour_props = 0;
    std::string section;
    while (std::getline(in, line)) {
      const auto w = words(lower(line));
      if (w.empty() || w[0] == "comment" || w[0] == "obj_info") continue;
      if (w[0] == "format") {
        if (w.size() < 2 || w[1] != "binary_little_endian")
          throw std::runtime_error("only binary_little_endian ply files are supported");
        le = true;
      } else if (w[0] == "element" && w.size() >= 3) {
        section = w[1];
        if (section == "vertex") { nv = std::stoll(w[2]); have_vertex = true; }
        else if (section == "face") { nf = std::stoll(w[2]); have_face = true; }
        else throw std::runtime_error("unsupported ply element: " + section);
      } else if (w[0] == "property") {
        if (section == "vertex") {
          if (w.size() < 3) throw std::runtime_error("bad vertex property");
          if (w[1] == "double" || w[1] == "float64") dbl = true;
          else if (w[1] != "float" && w[1] != "float32") throw std::runtime_error("Unknown vertices type.");
        } else if (section == "face") {
          if (w.size() >= 2 && w[1] == "list") continue;
          if (w.size() >= 3 && (w[2] == "red" || w[2] == "green" || w[2] == "blue")) ++face_colour_props;
          else throw std::runtime_error("unsupported face property: " + line);
        }
      } else if (w[0] == "end_header") {
        break;
      } else {
        throw std::runtime_error("unknown ply header line: " + line);
      }
    }
    if (!le) throw std::runtime_error("Unknown storage format, but we only support `binary_little_endian`.");
    if (!have_vertex || !have_face) throw std::runtime_error("ply file needs vertex and face elements");
    if (face_colour_props != 0 && face_colour_props != 3) throw std::runtime_error("Incomplete colors.");
    clear();
    v_.resize((size_t)nv);
    for (long long i = 0; i < nv; ++i) {
      if (dbl) {
        double p[3];
        in.read((char*)p, sizeof(p));
        v_[i] = {(float)p[0], (float)p[1], (float)p[2]};
      } else {
        float p[3];
        in.read((char*)p, sizeof(p));
        v_[i] = {p[0], p[1], p[2]};
      }
    }
    f_.resize((size_t)nf);
    for (long long i = 0; i < nf; ++i) {
      uint8_t n = 0;
      in.read((char*)&n, 1);
      f_[i].resize(n);
      if (n) in.read((char*)f_[i].data(), (std::streamsize)n * 4);
      if (face_colour_props == 3) {
        std::array<uint8_t, 3> c{};
        in.read((char*)c.data(), 3);
        c_.push_back(c);
      }
    }
    if (!in) throw std::runtime_error("ply file ends early: " + path);
  }

  void load_off(const std::string& path) {
    std::ifstream in(path);
    if (!in) throw std::runtime_error("cannot open " + path);
    std::string line;
    std::getline(in, line);
    if (lower(words(line).empty() ? "" : words(line)[0]) != "off")
      throw std::runtime_error("It should be an ascii off file!");
    auto next = [&]() {
      while (std::getline(in, line)) {
        const auto w = words(line);
        if (!w.empty() && w[0][0] != '#') return w;
      }
      throw std::runtime_error("off file ends early: " + path);
    };
    const auto counts = next();
    if (counts.size() < 2) throw std::runtime_error("bad off counts line");
    const long long nv = std::stoll(counts[0]), nf = std::stoll(counts[1]);
    clear();
    for (long long i = 0; i < nv; ++i) {
      const auto w = next();
      if (w.size() < 3) throw std::runtime_error("bad off vertex line");
      v_.push_back({std::stof(w[0]), std::stof(w[1]), std::stof(w[2])});
    }
    for (long long i = 0; i < nf; ++i) {
      const auto w = next();
      const size_t n = (size_t)std::stoll(w[0]);
      if (w.size() < 1 + n) throw std::runtime_error("bad off face line");
      std::vector<int> f(n);
      for (size_t j = 0; j < n; ++j) f[j] = std::stoi(w[1 + j]);
      f_.push_back(std::move(f));
      if (w.size() >= 1 + n + 3) {
        std::array<uint8_t, 3> c{};
        for (int q = 0; q < 3; ++q) {
          const float x = std::stof(w[1 + n + q]);
          // colours written as integers 0..255, or as 0..1 floats
          const bool unit = w[1 + n + q].find('.') != std::string::npos && x <= 1.0f;
          c[q] = (uint8_t)std::lround(std::min(255.0f, std::max(0.0f, unit ? x * 255.0f : x)));
        }
        c_.push_back(c);
      }
    }
  }

  void save_ply(const std::string& path) const {
    std::ofstream o(path, std::ios::binary);
    if (!o) throw std::runtime_error("cannot write " + path);
    o << "ply\nformat binary_little_endian 1.0\n";
    o << "element vertex " << v_.size() << "\n";
    o << "property float x\nproperty float y\nproperty float z\n";
    o << "element face " << f_.size() << "\n";
    o << "property list uchar int vertex_index\n";
    if (!c_.empty()) o << "property uchar red\nproperty uchar green\nproperty uchar blue\n";
    o << "end_header\n";
    for (const auto& p : v_) o.write((const char*)p.data(), 12);
    for (size_t i = 0; i < f_.size(); ++i) {
      if (f_[i].size() > 255) throw std::runtime_error("a face has more than 255 vertices");
      const uint8_t n = (uint8_t)f_[i].size();
      o.write((const char*)&n, 1);
      o.write((const char*)f_[i].data(), (std::streamsize)n * 4);
      if (!c_.empty()) o.write((const char*)c_[i].data(), 3);
    }
    if (!o) throw std::runtime_error("write failed: " + path);
  }

  void save_off(const std::string& path) const {
    std::ofstream o(path);
    if (!o) throw std::runtime_error("cannot write " + path);
    o.precision(9);
    o << "OFF\n" << v_.size() << " " << f_.size() << " 0\n";
    for (const auto& p : v_) o << p[0] << " " << p[1] << " " << p[2] << "\n";
    for (size_t i = 0; i < f_.size(); ++i) {
      o << f_[i].size();
      for (int x : f_[i]) o << " " << x;
      if (!c_.empty()) o << " " << (int)c_[i][0] << " " << (int)c_[i][1] << " " << (int)c_[i][2];
      o << "\n";
    }
    if (!o) throw std::runtime_error("write failed: " + path);
  }
};

}  // namespace

PYBIND11_MODULE(_polylib, m) {
  m.doc() = "Polygon-mesh file tools: PLY/OFF load and save, fan triangulation, face counts, vertex welding";
  py::class_<PolyMesh>(m, "PolyMesh")
      .def(py::init<>())
      .def(py::init<const std::string&>(), py::arg("load_path"))
      .def(py::init<const std::vector<std::vector<float>>&, const std::vector<std::vector<int>>&,
                    const std::vector<std::vector<int>>&>(),
           py::arg("vertices"), py::arg("faces"), py::arg("colors"))
      .def("clear", &PolyMesh::clear)
      .def("load", &PolyMesh::load, py::arg("load_path"))
      .def("save", &PolyMesh::save, py::arg("save_path"))
      .def("vertices", &PolyMesh::vertices)
      .def("faces", &PolyMesh::faces)
      .def("colors", &PolyMesh::colors)
      .def("set_vertices", &PolyMesh::set_vertices, py::arg("vertices"))
      .def("set_faces", &PolyMesh::set_faces, py::arg("faces"))
      .def("set_colors", &PolyMesh::set_colors, py::arg("colors"))
      .def("poly2tri", &PolyMesh::poly2tri)
      .def("is_polymesh", &PolyMesh::is_polymesh)
      .def("num_polyfaces", &PolyMesh::num_polyfaces)
      .def("num_trifaces", &PolyMesh::num_trifaces)
      .def("weld", &PolyMesh::weld, py::arg("tol"));
}
