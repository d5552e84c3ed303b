// pybind11 module with the reference's Python surface (lib/bindings.cpp:14-70):
// VecArray3d, VecArray3i, Mesh, MeshList, Config, make_vecarray3i, set_seed,
// config (a reference to the process-global Config), ProcessResult, process.
// Extra names with no counterpart in the reference: `engine` (EngineOptions),
// `last_profile` and a few host geometry services used by the tests.
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>
#include <pybind11/stl_bind.h>

#include "mesh.hpp"
#include "proposal.hpp"

namespace py = pybind11;

PYBIND11_MAKE_OPAQUE(std::vector<std::array<double, 3>>);
PYBIND11_MAKE_OPAQUE(std::vector<std::array<int, 3>>);

PYBIND11_MODULE(_visacd, m) {
  py::bind_vector<std::vector<std::array<double, 3>>>(m, "VecArray3d");
  py::bind_vector<std::vector<std::array<int, 3>>>(m, "VecArray3i");

  py::class_<vhost::Mesh>(m, "Mesh")
      .def(py::init<>())
      .def_readwrite("vertices", &vhost::Mesh::vertices)
      .def_readwrite("triangles", &vhost::Mesh::triangles);

  py::bind_vector<vhost::MeshList>(m, "MeshList");

  py::class_<vhost::Config>(m, "Config")
      .def(py::init<>())
      .def_readwrite("return_parts", &vhost::Config::return_parts)
      .def_readwrite("score_mode", &vhost::Config::score_mode)
      .def_readwrite("flat_surface_min_area", &vhost::Config::flat_surface_min_area)
      .def_readwrite("use_flat_surfaces", &vhost::Config::use_flat_surfaces)
      .def_readwrite("flat_surface_k", &vhost::Config::flat_surface_k)
      .def_readwrite("use_merging", &vhost::Config::use_merging);

  m.def("make_vecarray3i", [](py::array_t<int, py::array::c_style | py::array::forcecast> a) {
    if (a.ndim() != 2 || a.shape(1) != 3) throw std::invalid_argument("expected an (X, 3) array");
    std::vector<std::array<int, 3>> out((size_t)a.shape(0));
    auto r = a.unchecked<2>();
    for (py::ssize_t i = 0; i < a.shape(0); ++i) out[i] = {r(i, 0), r(i, 1), r(i, 2)};
    return out;
  });

  m.def("set_seed", &vhost::set_seed, py::arg("seed"));
  m.attr("config") = py::cast(&vhost::config, py::return_value_policy::reference);

  py::class_<vhost::ProcessResult>(m, "ProcessResult")
      .def_readwrite("parts", &vhost::ProcessResult::parts)
      .def_readwrite("concavity", &vhost::ProcessResult::concavity)
      .def_readwrite("num_parts", &vhost::ProcessResult::num_parts);

  m.def("process", &vhost::process, py::arg("mesh"), py::arg("concavity"), py::arg("num_parts"));
  m.def("last_profile", [] {
    py::dict d;
    for (auto &kv : vhost::last_profile.seconds)
      d[py::str(kv.first)] = py::make_tuple(kv.second, vhost::last_profile.calls[kv.first]);
    for (auto &kv : vhost::last_profile.calls)  // counters without a time
      if (!vhost::last_profile.seconds.count(kv.first)) d[py::str(kv.first)] = kv.second;
    return d;
  });

  // ---- the engine's own knobs (no counterpart in the reference) and host services
  py::class_<vhost::EngineOptions>(m, "EngineOptions")
      .def(py::init<>())
      .def_readwrite("device", &vhost::EngineOptions::device)
      .def_readwrite("max_points", &vhost::EngineOptions::max_points)
      .def_readwrite("num_planes", &vhost::EngineOptions::num_planes)
      .def_readwrite("planes", &vhost::EngineOptions::planes)
      .def_readwrite("hausdorff_resolution", &vhost::EngineOptions::hausdorff_resolution)
      .def_readwrite("sample_seed", &vhost::EngineOptions::sample_seed)
      .def_readwrite("gpu_distance", &vhost::EngineOptions::gpu_distance)
      .def_readwrite("cap_rule", &vhost::EngineOptions::cap_rule)
      .def_readwrite("visibility", &vhost::EngineOptions::visibility)
      .def_readwrite("remesh", &vhost::EngineOptions::remesh)
      .def_readwrite("mode", &vhost::EngineOptions::mode)
      .def_readwrite("batch_width", &vhost::EngineOptions::batch_width)
      .def_readwrite("eager_visibility", &vhost::EngineOptions::eager_visibility)
      .def_readwrite("front_max", &vhost::EngineOptions::front_max)
      .def_readwrite("proposal", &vhost::EngineOptions::proposal)
      .def_readwrite("proposal_weights", &vhost::EngineOptions::proposal_weights)
      .def_readwrite("proposal_top", &vhost::EngineOptions::proposal_top)
      .def_readwrite("proposal_audit", &vhost::EngineOptions::proposal_audit)
      .def_readwrite("proposal_audit_ratio", &vhost::EngineOptions::proposal_audit_ratio)
      .def_readwrite("planar_caps", &vhost::EngineOptions::planar_caps)
      .def_readwrite("verbose", &vhost::EngineOptions::verbose);
  m.attr("engine") = py::cast(&vhost::engine, py::return_value_policy::reference);

  m.def("convex_hull", [](const vhost::Mesh &mesh) { return vhost::convex_hull(mesh.vertices); });
  m.def("mesh_volume", [](const vhost::Mesh &mesh) { return vhost::mesh_volume(mesh); });
  m.def("surface_area", [](const vhost::Mesh &mesh) { return vhost::surface_area(mesh); });
  m.def("separate_disjoint", [](const vhost::Mesh &mesh) { return vhost::separate_disjoint(mesh); });
  m.def("clip", [](const vhost::Mesh &mesh, std::array<double, 4> p) {
    std::vector<int> mp, mn;
    vhost::MeshList parts = vhost::clip(mesh, vhost::Plane(p[0], p[1], p[2], p[3]), mp, mn);
    std::vector<std::vector<bool>> flags;
    for (auto &q : parts) flags.push_back(q.is_new);
    return py::make_tuple(parts, mp, mn, flags);
  });
  m.def("concavity", [](const vhost::Mesh &part, const vhost::Mesh &hull, int res, uint32_t seed) {
    return vhost::concavity(part, hull, res, seed);
  });
  // inputs of the learned plane proposal, as the driver computes them
  m.def("proposal_inputs",
        [](py::array_t<double, py::array::c_style | py::array::forcecast> verts,
           py::array_t<uint32_t, py::array::c_style | py::array::forcecast> edges,
           py::array_t<double, py::array::c_style | py::array::forcecast> planes, uint64_t seed) {
          if (verts.ndim() != 2 || verts.shape(1) != 3) throw std::invalid_argument("verts must be (n, 3)");
          if (edges.ndim() != 2 || edges.shape(1) != 2) throw std::invalid_argument("edges must be (E, 2)");
          if (planes.ndim() != 2 || planes.shape(1) != 4) throw std::invalid_argument("planes must be (P, 4)");
          const size_t n = (size_t)verts.shape(0), E = (size_t)edges.shape(0), P = (size_t)planes.shape(0);
          std::vector<vhost::Vec3> v(n);
          const double *pv = verts.data();
          for (size_t i = 0; i < n; ++i) v[i] = {pv[3 * i], pv[3 * i + 1], pv[3 * i + 2]};
          std::vector<std::pair<unsigned, unsigned>> e(E);
          const uint32_t *pe = edges.data();
          for (size_t i = 0; i < E; ++i) {
            if (pe[2 * i] >= n || pe[2 * i + 1] >= n) throw std::out_of_range("edge endpoint out of range");
            e[i] = {pe[2 * i], pe[2 * i + 1]};
          }
          std::vector<vhost::Plane> pl(P);
          const double *pp = planes.data();
          for (size_t i = 0; i < P; ++i) pl[i] = vhost::Plane(pp[4 * i], pp[4 * i + 1], pp[4 * i + 2], pp[4 * i + 3]);
          const vhost::PartFrame fr = vhost::part_frame(v);
          py::array_t<float> feat({(py::ssize_t)vhost::kProposalEdges, (py::ssize_t)vhost::kEdgeFeatures});
          py::array_t<float> q({(py::ssize_t)P, (py::ssize_t)4});
          vhost::edge_features(v, e, seed, fr, feat.mutable_data());
          vhost::plane_inputs(pl, P, fr, q.mutable_data());
          return py::make_tuple(feat, q, py::make_tuple(fr.c[0], fr.c[1], fr.c[2], fr.s));
        },
        py::arg("verts"), py::arg("edges"), py::arg("planes"), py::arg("seed"));
  m.def("bisector_planes",
        [](py::array_t<double, py::array::c_style | py::array::forcecast> verts,
           py::array_t<uint32_t, py::array::c_style | py::array::forcecast> edges, int num_planes,
           uint32_t seed) {
          if (verts.ndim() != 2 || verts.shape(1) != 3) throw std::invalid_argument("verts must be (n, 3)");
          if (edges.ndim() != 2 || edges.shape(1) != 2) throw std::invalid_argument("edges must be (E, 2)");
          vhost::Mesh m;
          const size_t n = (size_t)verts.shape(0);
          m.vertices.resize(n);
          for (size_t i = 0; i < n; ++i)
            m.vertices[i] = {verts.data()[3 * i], verts.data()[3 * i + 1], verts.data()[3 * i + 2]};
          for (py::ssize_t e = 0; e < edges.shape(0); ++e) {
            const uint32_t a = edges.data()[2 * e], b = edges.data()[2 * e + 1];
            if (a >= n || b >= n) throw std::out_of_range("edge endpoint out of range");
            m.intersecting_edges.push_back({a, b});
          }
          std::vector<std::array<double, 4>> out;
          for (auto &p : vhost::bisector_planes_seeded(m, num_planes, seed)) out.push_back({p.a, p.b, p.c, p.d});
          return out;
        },
        py::arg("verts"), py::arg("edges"), py::arg("num_planes"), py::arg("seed"));
  m.def("select_plane", &vhost::select_plane, py::arg("score_q32"), py::arg("ss_offset"));
  // the driver's split of a part given as arrays; None when the clip fails,
  // else a list of (vertices, triangles, is_new, edges, points, cut_planes, seed)
  m.def("split_part",
        [](py::array_t<double, py::array::c_style | py::array::forcecast> V,
           py::array_t<int, py::array::c_style | py::array::forcecast> T,
           py::array_t<uint32_t, py::array::c_style | py::array::forcecast> E,
           py::array_t<int, py::array::c_style | py::array::forcecast> points,
           py::array_t<double, py::array::c_style | py::array::forcecast> cuts, uint64_t seed,
           std::array<double, 4> plane) -> py::object {
          vhost::Mesh m;
          const size_t nv = (size_t)V.shape(0);
          for (size_t i = 0; i < nv; ++i) m.vertices.push_back({V.data()[3 * i], V.data()[3 * i + 1], V.data()[3 * i + 2]});
          for (py::ssize_t t = 0; t < T.shape(0); ++t)
            m.triangles.push_back({T.data()[3 * t], T.data()[3 * t + 1], T.data()[3 * t + 2]});
          for (py::ssize_t e = 0; e < E.shape(0); ++e) {
            if (E.data()[2 * e] >= nv || E.data()[2 * e + 1] >= nv) throw std::out_of_range("edge endpoint out of range");
            m.intersecting_edges.push_back({E.data()[2 * e], E.data()[2 * e + 1]});
          }
          for (py::ssize_t i = 0; i < points.shape(0); ++i) m.points.push_back(points.data()[i]);
          for (py::ssize_t c = 0; c < cuts.shape(0); ++c)
            m.cut_planes.emplace_back(cuts.data()[4 * c], cuts.data()[4 * c + 1], cuts.data()[4 * c + 2],
                                      cuts.data()[4 * c + 3]);
          vhost::MeshList kids;
          std::vector<uint64_t> ks;
          if (!vhost::split_part_seeded(m, seed, vhost::Plane(plane[0], plane[1], plane[2], plane[3]), kids, ks))
            return py::none();
          py::list out;
          for (size_t k = 0; k < kids.size(); ++k) {
            const vhost::Mesh &q = kids[k];
            py::array_t<double> kv({(py::ssize_t)q.vertices.size(), (py::ssize_t)3});
            for (size_t i = 0; i < q.vertices.size(); ++i)
              for (int c = 0; c < 3; ++c) kv.mutable_data()[3 * i + c] = q.vertices[i][c];
            py::array_t<int> kt({(py::ssize_t)q.triangles.size(), (py::ssize_t)3});
            for (size_t t = 0; t < q.triangles.size(); ++t)
              for (int c = 0; c < 3; ++c) kt.mutable_data()[3 * t + c] = q.triangles[t][c];
            py::array_t<bool> kn((py::ssize_t)q.vertices.size());
            for (size_t i = 0; i < q.vertices.size(); ++i) kn.mutable_data()[i] = i < q.is_new.size() && q.is_new[i];
            py::array_t<uint32_t> ke({(py::ssize_t)q.intersecting_edges.size(), (py::ssize_t)2});
            for (size_t e = 0; e < q.intersecting_edges.size(); ++e)
              ke.mutable_data()[2 * e] = q.intersecting_edges[e].first, ke.mutable_data()[2 * e + 1] = q.intersecting_edges[e].second;
            py::array_t<int> kp((py::ssize_t)q.points.size());
            for (size_t i = 0; i < q.points.size(); ++i) kp.mutable_data()[i] = q.points[i];
            py::array_t<double> kc({(py::ssize_t)q.cut_planes.size(), (py::ssize_t)4});
            for (size_t c = 0; c < q.cut_planes.size(); ++c) {
              const vhost::Plane &pl = q.cut_planes[c];
              kc.mutable_data()[4 * c] = pl.a, kc.mutable_data()[4 * c + 1] = pl.b;
              kc.mutable_data()[4 * c + 2] = pl.c, kc.mutable_data()[4 * c + 3] = pl.d;
            }
            out.append(py::make_tuple(kv, kt, kn, ke, kp, kc, ks[k]));
          }
          return out;
        },
        py::arg("vertices"), py::arg("triangles"), py::arg("edges"), py::arg("points"),
        py::arg("cut_planes"), py::arg("seed"), py::arg("plane"));
  m.def("sphere_draws", [](int num_planes, uint32_t seed) {
    std::vector<double> n, f;
    vhost::sphere_draws_seeded(num_planes, seed, n, f);
    py::array_t<double> an({(py::ssize_t)f.size(), (py::ssize_t)3});
    std::copy(n.begin(), n.end(), an.mutable_data());
    py::array_t<double> af((py::ssize_t)f.size());
    std::copy(f.begin(), f.end(), af.mutable_data());
    return py::make_tuple(an, af);
  });
  m.def("hausdorff_boundary", [](const vhost::Mesh &a, const vhost::Mesh &b, int res, uint32_t seed) {
    return vhost::hausdorff_boundary(a, b, res, seed);
  });
  m.def("flat_surface_planes", [](const vhost::Mesh &mesh, double min_area) {
    std::vector<std::array<double, 4>> out;
    for (auto &p : vhost::flat_surface_planes(mesh, min_area)) out.push_back({p.a, p.b, p.c, p.d});
    return out;
  });
}
