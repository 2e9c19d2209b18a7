// Python module `cubff` with the reference's bindings (cubff_py.cc of
// paradigms-of-intelligence/cubff): SimulationParams, SimulationState,
// LanguageInterface, GetLanguage, ResetColors and kSelfrepThreshold, with the
// soup and the pairing exposed as buffer-protocol vectors (VectorUint8,
// VectorUint32) so numpy views them without a copy.
#include <pybind11/functional.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>
#include <pybind11/stl_bind.h>

#include <stdint.h>
#include <stdio.h>

#include <vector>

#include "cubff.h"

namespace py = pybind11;

PYBIND11_MAKE_OPAQUE(std::vector<uint8_t>);
PYBIND11_MAKE_OPAQUE(std::vector<uint32_t>);

namespace {

template <typename C, typename... F>
void readwrite(py::class_<C> &cls, F... fields) {
  (cls.def_readwrite(fields.first, fields.second), ...);
}

}  // namespace

PYBIND11_MODULE(cubff, m) {
  m.doc() = "cubff python module (MI355X build: epochs run on the GPU)";

  py::class_<SimulationParams> params(m, "SimulationParams");
  params.def(py::init<>());
  readwrite(params, std::make_pair("num_programs", &SimulationParams::num_programs),
            std::make_pair("seed", &SimulationParams::seed),
            std::make_pair("mutation_prob", &SimulationParams::mutation_prob),
            std::make_pair("reset_interval", &SimulationParams::reset_interval),
            std::make_pair("load_from", &SimulationParams::load_from),
            std::make_pair("save_to", &SimulationParams::save_to),
            std::make_pair("callback_interval", &SimulationParams::callback_interval),
            std::make_pair("save_interval", &SimulationParams::save_interval),
            std::make_pair("permute_programs", &SimulationParams::permute_programs),
            std::make_pair("fixed_shuffle", &SimulationParams::fixed_shuffle),
            std::make_pair("zero_init", &SimulationParams::zero_init),
            std::make_pair("allowed_interactions", &SimulationParams::allowed_interactions),
            std::make_pair("eval_selfrep", &SimulationParams::eval_selfrep));

  py::bind_vector<std::vector<uint8_t>>(m, "VectorUint8", py::buffer_protocol());
  py::bind_vector<std::vector<uint32_t>>(m, "VectorUint32", py::buffer_protocol());

  py::class_<SimulationState>(m, "SimulationState")
      .def_readonly("soup", &SimulationState::soup)
      .def_readonly("shuffle_idx", &SimulationState::shuffle_idx)
      .def_readonly("elapsed_s", &SimulationState::elapsed_s)
      .def_readonly("total_ops", &SimulationState::total_ops)
      .def_readonly("mops_s", &SimulationState::mops_s)
      .def_readonly("epoch", &SimulationState::epoch)
      .def_readonly("ops_per_run", &SimulationState::ops_per_run)
      .def_readonly("brotli_size", &SimulationState::brotli_size)
      .def_readonly("brotli_bpb", &SimulationState::brotli_bpb)
      .def_readonly("bytes_per_prog", &SimulationState::bytes_per_prog)
      .def_readonly("h0", &SimulationState::h0)
      .def_readonly("higher_entropy", &SimulationState::higher_entropy)
      .def_readonly("frequent_bytes", &SimulationState::frequent_bytes)
      .def_readonly("uncommon_bytes", &SimulationState::uncommon_bytes)
      .def_readonly("replication_per_prog", &SimulationState::replication_per_prog);

  py::class_<LanguageInterface>(m, "LanguageInterface")
      .def("PrintProgram",
           [](const LanguageInterface *lang, size_t pc_pos, const std::vector<uint8_t> &program,
              const std::vector<size_t> &separators) {
             lang->PrintProgram(pc_pos, program.data(), program.size(), separators.data(),
                                separators.size());
           })
      .def("RunSimulation", &LanguageInterface::RunSimulation)
      .def("RunSingleProgram", &LanguageInterface::RunSingleProgram)
      .def("RunSingleParsedProgram", &LanguageInterface::RunSingleParsedProgram)
      .def("EvalSelfrep", &LanguageInterface::EvalSelfrep)
      .def("EvalParsedSelfrep", &LanguageInterface::EvalParsedSelfrep)
      .def("Parse", &LanguageInterface::Parse);

  m.def("GetLanguage", &GetLanguage, py::return_value_policy::reference);
  m.def("ResetColors", [] {
    fputs(ResetColors(), stdout);
    fflush(stdout);
  });
  m.attr("kSelfrepThreshold") = kSelfrepThreshold;
}
