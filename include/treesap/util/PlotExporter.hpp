// PlotExporter.hpp -- writes the state every `export_interval` steps as a
// single-level plotfile named path + step (zero-padded to 5 digits), like
// util/PlotExporter.hpp:29-67 does through amrex::WriteSingleLevelPlotfile.
// The layout follows AMReX's documented plotfile format: a HyperCLaw-V1.1 text
// header (with the physical extent of every box of the level), a
// Level_0/Cell_H FabArray header (the box array, one FabOnDisk entry per box
// with its byte offset, per-box component minima and maxima) and one binary
// file Level_0/Cell_D_00000 holding one FAB per box, each a text line with the
// native-order IEEE double descriptor, the box and the component count,
// followed by the box's values (x fastest, component slowest). The boxes are
// those of the integrator's level (one box unless block_size splits the
// domain); the state is copied from the device for each export.
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
    // `boxes`: the level's boxes (empty: the domain as one box); together they must tile the domain.
    PlotExporter(std::function<void(std::vector<amrex::Real> &)> fetch, const amrex::Geometry &geom,
                 amrex::Vector<std::string> comp_names, fs::path path, int export_interval,
                 std::vector<amrex::Box> boxes = {})
        : fetch_(std::move(fetch)), geom_(geom), names_(std::move(comp_names)), path_(std::move(path)),
          interval_(export_interval), boxes_(std::move(boxes)) {
        if (interval_ <= 0) throw std::invalid_argument("treesap: export interval must be positive");
        if (boxes_.empty()) boxes_.push_back(geom_.Domain());
        std::int64_t cells = 0;
        for (const auto &b : boxes_) {
            for (int d = 0; d < 3; ++d)
                if (b.smallEnd(d) < geom_.Domain().smallEnd(d) || b.bigEnd(d) > geom_.Domain().bigEnd(d))
                    throw std::invalid_argument("treesap: plotfile box outside the domain");
            cells += b.numPts();
        }
        if (cells != geom_.Domain().numPts()) throw std::invalid_argument("treesap: plotfile boxes do not tile the domain");
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
        const int nbox = (int)boxes_.size();

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
        h << 0 << ' ' << nbox << ' ' << time << '\n' << step << '\n';
        for (const auto &b : boxes_)
            for (int d = 0; d < 3; ++d)
                h << geom_.ProbLo(d) + b.smallEnd(d) * geom_.CellSize(d) << ' '
                  << geom_.ProbLo(d) + (b.bigEnd(d) + 1) * geom_.CellSize(d) << '\n';
        h << "Level_0/Cell\n";
        if (!h) throw std::runtime_error("treesap: cannot write " + dir + "/Header");

        // FABs, one per box, and their offsets in Cell_D_00000
        std::ofstream fab(fs::path(dir) / "Level_0" / "Cell_D_00000", std::ios::binary);
        std::vector<std::int64_t> offset(nbox);
        std::vector<std::vector<double>> lo(nbox, std::vector<double>(nc)), hi(nbox, std::vector<double>(nc));
        std::vector<double> buf;
        const std::int64_t nx = dom.length(0), ny = dom.length(1), nz = dom.length(2);
        for (int ib = 0; ib < nbox; ++ib) {
            const amrex::Box &b = boxes_[ib];
            const std::int64_t bc = b.numPts();
            buf.resize((std::size_t)(bc * nc));
            std::size_t q = 0;
            for (int c = 0; c < nc; ++c)
                for (int k = b.smallEnd(2); k <= b.bigEnd(2); ++k)
                    for (int j = b.smallEnd(1); j <= b.bigEnd(1); ++j) {
                        const std::int64_t row =
                            ((c * nz + (k - dom.smallEnd(2))) * ny + (j - dom.smallEnd(1))) * nx - dom.smallEnd(0);
                        for (int i = b.smallEnd(0); i <= b.bigEnd(0); ++i) buf[q++] = state[(std::size_t)(row + i)];
                    }
            for (int c = 0; c < nc; ++c) {
                const auto first = buf.begin() + c * bc;
                lo[ib][c] = *std::min_element(first, first + bc);
                hi[ib][c] = *std::max_element(first, first + bc);
            }
            offset[ib] = (std::int64_t)fab.tellp();
            fab << "FAB ((8, (64 11 52 0 1 12 0 1023)),(8, (8 7 6 5 4 3 2 1)))" << box_text(b) << ' ' << nc << '\n';
            fab.write(reinterpret_cast<const char *>(buf.data()), (std::streamsize)(buf.size() * sizeof(double)));
        }
        if (!fab) throw std::runtime_error("treesap: cannot write " + dir + "/Level_0/Cell_D_00000");

        std::ofstream ch(fs::path(dir) / "Level_0" / "Cell_H");
        ch << 1 << '\n' << 1 << '\n' << nc << '\n' << 0 << '\n';
        ch << '(' << nbox << " 0\n";
        for (const auto &b : boxes_) ch << box_text(b) << '\n';
        ch << ")\n";
        ch << nbox << '\n';
        for (int ib = 0; ib < nbox; ++ib) ch << "FabOnDisk: Cell_D_00000 " << offset[ib] << '\n';
        ch << '\n';
        ch << std::setprecision(16) << std::scientific;
        ch << nbox << ',' << nc << '\n';
        for (int ib = 0; ib < nbox; ++ib) {
            for (double v : lo[ib]) ch << v << ',';
            ch << '\n';
        }
        ch << '\n' << nbox << ',' << nc << '\n';
        for (int ib = 0; ib < nbox; ++ib) {
            for (double v : hi[ib]) ch << v << ',';
            ch << '\n';
        }
        if (!ch) throw std::runtime_error("treesap: cannot write " + dir + "/Level_0/Cell_H");
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
    std::vector<amrex::Box> boxes_;
};

} // namespace treesap
