// pybind11 module pyPanSim: the reference's Python binding of
// SimulatorInterface (matlab/mexPanSim.cpp:57-127, module definition
// :207-219), over the host-loop C ABI (include/pansim_sim.h), so that a
// calibration script written for the reference runs unchanged:
//
//     import pyPanSim
//     s = pyPanSim.SimulatorInterface()
//     s.initSimulation(["panSim", "-w", "4", "--quarantinePolicy", "3", ...])
//     stats = s.runForDay(["TPdef", "PLNONE", "CFNONE", "SONONE", "QU3", "MA1.0"])
//
// Behaviour kept from the reference:
//  * initSimulation receives argv with the program name first (the option
//    parser skips argv[0]);
//  * runForDay with an empty flag list runs nothing and returns the previous
//    day's statistics (the reference calls Simulation::runForDay only for a
//    non-empty list, mexPanSim.cpp:113-114), an empty list before the first
//    day;
//  * runForDay before initSimulation prints "Cannot run uninitialized
//    simulation... Initialize it first." and returns an empty list;
//  * an error inside runForDay is written to stderr and the call returns an
//    empty list (mexPanSim.cpp:119-122);
//  * the object's destruction finalizes the simulation (output files and the
//    --diags summary, mexPanSim.cpp:70-73); close() does it at a chosen
//    point.
// Differences: a failed initSimulation raises RuntimeError with the
// initialisation error (the reference prints it and keeps an unusable
// simulation), and input paths that do not exist relative to the working
// directory are looked up in the input files shipped with this package
// (--dataDir, unless the caller gives one).
#include <dlfcn.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <algorithm>
#include <cstdio>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../../include/pansim_sim.h"

namespace py = pybind11;

namespace {

// <directory of this module>/../data: the package's input files
std::string shipped_data_dir() {
    Dl_info info;
    if (!dladdr(reinterpret_cast<void*>(&shipped_data_dir), &info) || !info.dli_fname) return std::string();
    std::string p = info.dli_fname;
    const size_t s = p.rfind('/');
    return (s == std::string::npos ? std::string(".") : p.substr(0, s)) + "/../data";
}

class SimulatorInterface {
    pansim_sim* sim_ = nullptr;
    std::vector<unsigned> stats_;

public:
    SimulatorInterface() = default;
    SimulatorInterface(const SimulatorInterface&) = delete;
    SimulatorInterface& operator=(const SimulatorInterface&) = delete;
    ~SimulatorInterface() { close(); }

    void initSimulation(const std::vector<std::string>& options) {
        close();
        std::vector<std::string> args(options.begin() + (options.empty() ? 0 : 1), options.end());
        if (std::find(args.begin(), args.end(), "--dataDir") == args.end()) {
            const std::string d = shipped_data_dir();
            if (!d.empty()) {
                args.push_back("--dataDir");
                args.push_back(d);
            }
        }
        std::vector<const char*> argv;
        for (const auto& a : args) argv.push_back(a.c_str());
        int rc;
        {
            py::gil_scoped_release nogil;
            rc = pansim_sim_create(&sim_, (int)argv.size(), argv.data(), 0);
        }
        if (rc != 0) {
            sim_ = nullptr;
            throw std::runtime_error(pansim_sim_create_error());
        }
        stats_.clear();
    }

    std::vector<unsigned> runForDay(const std::vector<std::string>& options) {
        if (!sim_) {
            std::printf("Cannot run uninitialized simulation... Initialize it first.");
            std::fflush(stdout);
            return std::vector<unsigned>();
        }
        if (options.empty()) return stats_;
        std::vector<const char*> flags;
        for (const auto& f : options) flags.push_back(f.c_str());
        std::vector<uint32_t> out(4096);
        size_t len = 0;
        int rc;
        {
            py::gil_scoped_release nogil;
            rc = pansim_sim_run_for_day(sim_, (int)flags.size(), flags.data(), out.data(), out.size(), &len);
        }
        if (rc != 0) {
            std::cerr << pansim_sim_error(sim_);
            std::cerr.flush();
            return std::vector<unsigned>();
        }
        stats_.assign(out.begin(), out.begin() + (long)std::min(len, out.size()));
        return stats_;
    }

    // finalize (output files, --diags summary) and release the device
    // memory; the object is uninitialized afterwards
    void close() {
        if (!sim_) return;
        pansim_sim* s = sim_;
        sim_ = nullptr;
        const int rc = pansim_sim_finalize(s);
        if (rc != 0) std::cerr << pansim_sim_error(s) << std::endl;
        pansim_sim_destroy(s);
    }
};

}  // namespace

PYBIND11_MODULE(pyPanSim, m) {
    m.doc() = "PanSim on MI355X: SimulatorInterface (initSimulation / runForDay)";
    py::class_<SimulatorInterface>(m, "SimulatorInterface")
        .def(py::init<>())
        .def("initSimulation", &SimulatorInterface::initSimulation, py::arg("options"))
        .def("runForDay", &SimulatorInterface::runForDay, py::arg("options"))
        .def("close", &SimulatorInterface::close);
}
