// pybind11 module diff_tritracer._C: the bindings of the reference's ext.cpp:11-27
// (FwdTraceResult / BwdTraceResult with .data, OptixTracer with two constructors,
// build_tri_accel, trace_forward, trace_backward), plus the tracer's options.
#include <pybind11/pybind11.h>
#include <torch/extension.h>

#include "tracer.h"

PYBIND11_MODULE(TORCH_EXTENSION_NAME, m) {
  pybind11::class_<FwdTraceResult>(m, "FwdTraceResult")
      .def(pybind11::init<>())
      .def_readonly("data", &FwdTraceResult::data)
      .def_readonly("hits", &FwdTraceResult::hits)
      .def_readonly("hit_counts", &FwdTraceResult::hit_counts)
      .def_readonly("hit_colors", &FwdTraceResult::hit_colors);

  pybind11::class_<BwdTraceResult>(m, "BwdTraceResult")
      .def(pybind11::init<>())
      .def_readonly("data", &BwdTraceResult::data);

  pybind11::class_<OptixTracer>(m, "OptixTracer")
      .def(pybind11::init<const std::string &>())
      .def(pybind11::init<const std::string &, const std::string &, const std::string &>())
      .def("build_tri_accel", &OptixTracer::build_tri_accel)
      .def("trace_forward", &OptixTracer::trace_fwd)
      .def("trace_backward", &OptixTracer::trace_bwd, pybind11::arg("ray_o"), pybind11::arg("ray_d"),
           pybind11::arg("background"), pybind11::arg("triangles_points"), pybind11::arg("sigma"),
           pybind11::arg("num_points_per_triangle"), pybind11::arg("cumsum_of_points_per_triangle"),
           pybind11::arg("colors"), pybind11::arg("others"), pybind11::arg("number_of_points"), pybind11::arg("shs"),
           pybind11::arg("degree"), pybind11::arg("opacities"), pybind11::arg("dL_dpixels"),
           pybind11::arg("dL_dothers"), pybind11::arg("scaling"), pybind11::arg("debug"),
           pybind11::arg("hits") = pybind11::none(), pybind11::arg("hit_counts") = pybind11::none(),
           pybind11::arg("hit_colors") = pybind11::none())
      .def("set_options", &OptixTracer::set_options, pybind11::arg("k") = 16, pybind11::arg("cutoff") = 1e-4,
           pybind11::arg("chunk_mode") = 0, pybind11::arg("hit_cache") = 64)
      .def("get_options", &OptixTracer::get_options)
      .def("set_store_hits", &OptixTracer::set_store_hits, pybind11::arg("store"))
      .def("set_rebuild_policy", &OptixTracer::set_rebuild_policy, pybind11::arg("max_ratio"))
      .def("build_stats", &OptixTracer::build_stats)
      .def_property_readonly("variant", &OptixTracer::variant_name);

  m.def("version", []() { return std::string(tt_version()); });
}
