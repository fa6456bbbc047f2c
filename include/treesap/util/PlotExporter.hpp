// PlotExporter.hpp -- writes the state every `export_interval` steps as a
// single-level, single-box plotfile named path + step (zero-padded to 5
// digits), like util/PlotExporter.hpp:29-67 does through
// amrex::WriteSingleLevelPlotfile. The layout follows AMReX's documented
// plotfile format (HyperCLaw-V1.1 text header, a Level_0/Cell_H FabArray
// header and one binary FAB file with a native-order IEEE double descriptor);
// the state is copied from the device for each export.
#pragma once

#include "../amrex_types.hpp"
#include "ScopedTimer.hpp"

#include <algorithm>
#include <cstdint>
#include <filesystem>
#include <fstream>
#include <functional>
#include <iomanip>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace treesap {

namespace fs = std::filesystem;

class PlotExporter {
  public:
    // `fetch` fills a host buffer with the valid state, a[((c*nz+k)*ny+j)*nx+i].
    PlotExporter(std::function<void(std::vector<amrex::Real> &)> fetch, const amrex::Geometry &geom,
                 amrex::Vector<std::string> comp_names, fs::path path, int export_interval)
        : fetch_(std::move(fetch)), geom_(geom), names_(std::move(comp_names)), path_(std::move(path)),
          interval_(export_interval) {
        if (interval_ <= 0) throw std::invalid_argument("treesap: export interval must be positive");
    }

    void exportIfNecessary(std::int64_t step, amrex::Real time) const {
        if (step % interval_ != 0) return;
        ScopedTimer timer("PlotExporter::exportIfNecessary()");
        std::vector<amrex::Real> state;
        fetch_(state);
        write_plotfile(plotfile_name(step), state, time, step);
    }

    std::string plotfile_name(std::int64_t step) const {
        std::ostringstream o;
        o << path_.string() << std::setfill('0') << std::setw(5) << step;
        return o.str();
    }

    void write_plotfile(const std::string &dir, const std::vector<amrex::Real> &state, amrex::Real time,
                        std::int64_t step) const {
        const amrex::Box &dom = geom_.Domain();
        const int nc = (int)names_.size();
        const std::int64_t cells = dom.numPts();
        if ((std::int64_t)state.size() != cells * nc)
            throw std::invalid_argument("treesap: state size does not match the geometry and component names");
        fs::create_directories(fs::path(dir) / "Level_0");

        std::ofstream h(fs::path(dir) / "Header");
        h << std::setprecision(17);
        h << "HyperCLaw-V1.1\n" << nc << '\n';
        for (const auto &n : names_) h << n << '\n';
        h << 3 << '\n' << time << '\n' << 0 << '\n';
        for (int d = 0; d < 3; ++d) h << geom_.ProbLo(d) << ' ';
        h << '\n';
        for (int d = 0; d < 3; ++d) h << geom_.ProbHi(d) << ' ';
        h << '\n' << '\n';
        h << box_text(dom) << ' ' << '\n';
        h << step << ' ' << '\n';
        for (int d = 0; d < 3; ++d) h << geom_.CellSize(d) << ' ';
        h << '\n' << geom_.Coord() << '\n' << 0 << '\n';
        h << 0 << ' ' << 1 << ' ' << time << '\n' << step << '\n';
        for (int d = 0; d < 3; ++d)
            h << geom_.ProbLo(d) + dom.smallEnd(d) * geom_.CellSize(d) << ' '
              << geom_.ProbLo(d) + (dom.bigEnd(d) + 1) * geom_.CellSize(d) << '\n';
        h << "Level_0/Cell\n";
        if (!h) throw std::runtime_error("treesap: cannot write " + dir + "/Header");

        std::ofstream ch(fs::path(dir) / "Level_0" / "Cell_H");
        ch << 1 << '\n' << 1 << '\n' << nc << '\n' << 0 << '\n';
        ch << "(1 0\n" << box_text(dom) << "\n)\n";
        ch << 1 << '\n' << "FabOnDisk: Cell_D_00000 0\n\n";
        std::vector<double> lo(nc), hi(nc);
        for (int c = 0; c < nc; ++c) {
            const auto b = state.begin() + c * cells;
            lo[c] = cells ? *std::min_element(b, b + cells) : 0.0;
            hi[c] = cells ? *std::max_element(b, b + cells) : 0.0;
        }
        ch << std::setprecision(16) << std::scientific;
        ch << 1 << ',' << nc << '\n';
        for (double v : lo) ch << v << ',';
        ch << "\n\n" << 1 << ',' << nc << '\n';
        for (double v : hi) ch << v << ',';
        ch << '\n';
        if (!ch) throw std::runtime_error("treesap: cannot write " + dir + "/Level_0/Cell_H");

        std::ofstream fab(fs::path(dir) / "Level_0" / "Cell_D_00000", std::ios::binary);
        fab << "FAB ((8, (64 11 52 0 1 12 0 1023)),(8, (8 7 6 5 4 3 2 1)))" << box_text(dom) << ' ' << nc << '\n';
        fab.write(reinterpret_cast<const char *>(state.data()), (std::streamsize)(state.size() * sizeof(double)));
        if (!fab) throw std::runtime_error("treesap: cannot write " + dir + "/Level_0/Cell_D_00000");
    }

  private:
    static std::string box_text(const amrex::Box &b) {
        std::ostringstream o;
        o << "((" << b.smallEnd(0) << ',' << b.smallEnd(1) << ',' << b.smallEnd(2) << ") (" << b.bigEnd(0) << ','
          << b.bigEnd(1) << ',' << b.bigEnd(2) << ") (0,0,0))";
        return o.str();
    }

    std::function<void(std::vector<amrex::Real> &)> fetch_;
    amrex::Geometry geom_;
    amrex::Vector<std::string> names_;
    fs::path path_;
    int interval_;
};

} // namespace treesap
