// openTM_pybind.cpp -- the pybind11 module `openTM` the reference documents
// (main.cpp:90-120, commented out there; usage in python/testpyd.py):
// enums InitWay and Model, runInstance, and class homo (wrapper_homo) with
// setDensity / setConfig / optimize. Built on the C++ drop-in header, so the
// engine underneath is the gfx950 library.
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include "../../include/opentm/opentm.hpp"

namespace py = pybind11;

PYBIND11_MODULE(openTM, m) {
  m.doc() = "thermal inverse homogenization on MI355X (gfx950)";
  py::enum_<cfg::InitWay>(m, "InitWay")
      .value("random", cfg::InitWay::random)
      .value("randcenter", cfg::InitWay::randcenter)
      .value("noise", cfg::InitWay::noise)
      .value("manual", cfg::InitWay::manual)
      .value("interp", cfg::InitWay::interp)
      .value("rep_randcenter", cfg::InitWay::rep_randcenter)
      .value("P", cfg::InitWay::P)
      .value("G", cfg::InitWay::G)
      .value("D", cfg::InitWay::D)
      .value("IWP", cfg::InitWay::IWP)
      .value("example", cfg::InitWay::example);
  py::enum_<cfg::Model>(m, "Model").value("mma", cfg::Model::mma).value("oc", cfg::Model::oc);
  m.def("runInstance", &runInstance, py::arg("reso"), py::arg("heat_ratios"), py::arg("target_ratio"),
        py::arg("initway") = cfg::InitWay::IWP, py::arg("model") = cfg::Model::mma,
        py::call_guard<py::gil_scoped_release>());
  py::class_<wrapper_homo>(m, "homo")
      .def(py::init<>())
      .def("setDensity", &wrapper_homo::setDensity, "set density by given rho")
      .def("setConfig", &wrapper_homo::setConfig, py::arg("reso"), py::arg("heat_ratios"), py::arg("target_ratio"),
           py::arg("model") = cfg::Model::oc)
      .def("optimize", &wrapper_homo::optimize, "optimization", py::call_guard<py::gil_scoped_release>())
      // not in the reference: iteration cap and initialiser seed of the next optimize()
      .def("setMaxIter", [](wrapper_homo& w, int n) { w.configuration().max_iter = n; })
      .def("setSeed", [](wrapper_homo& w, int s) { w.configuration().seed = s; });
}
